// output.h -- asynchronous .dat output for instances that keep stepping.
//
// Extension with no counterpart in the reference, whose writeFiles
// (gel.cu:543-553, fluid.cu:140-150) copies the fields to the host
// synchronously before its writer thread formats them.  Here a request
// snapshots the output fields device-to-device on the library stream (in
// order with the step kernels, microseconds), copies the snapshot to pinned
// host memory on a separate copy stream, and a writer thread formats the
// files once that copy has landed, while the instance goes on stepping.
// The files and their layout are the reference's (gel<id>{rn,rm,um,vm,wm,
// Fn,Veln}<t>.dat, gel<id>bodycenter.dat, Velb<t>.dat, Conc<t>.dat).
#pragma once

#include <memory>
#include <string>

class Gel;
class Fluid;

namespace glsm {
struct OutputState;
}

class OutputStream {
 public:
  // `writers` threads format files in parallel (requests for one object
  // stay in order: each waits for the previous one of that object)
  explicit OutputStream(int writers = 4);
  ~OutputStream();  // waits for every pending write
  OutputStream(const OutputStream&) = delete;
  OutputStream& operator=(const OutputStream&) = delete;

  // The files of output time `time` for the gel's / fluid's current state,
  // written into `dir` ("" = the working directory).  The next request for
  // the same object waits until this one has been written.
  void gel(Gel* g, int time, const std::string& dir);
  void fluid(Fluid* f, int time, const std::string& dir);
  // Forget the buffers kept for an object (after its last request).
  void release(const void* object);

  void flush();             // wait until every request has been written
  long pending() const;     // requests not yet written
  long completed() const;   // requests written
  std::string lastError();  // "" if every write succeeded

 private:
  std::unique_ptr<glsm::OutputState> st_;
};
