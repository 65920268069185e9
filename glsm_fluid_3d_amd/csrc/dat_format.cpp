// dat_format.cpp -- the reference's .dat text layout, written from host
// arrays.  Gel::recordData, Fluid::recordData and the asynchronous output
// stream all write through these functions.
//
//   Velb<t>.dat  one line per lattice cell (x fastest):
//                setw(9) << to_string(float) for ux, uy, uz, joined by six
//                spaces (fluid.cu:115-138)
//   Conc<t>.dat  one column, same format
//   gel<id>{rn,Fn,Veln}<t>.dat  real nodes 1..L (z, y, x loops), 3 columns
//   gel<id>rm<t>.dat            interior elements 1..L-1, 3 columns
//   gel<id>um<t>.dat            interior elements, to_string
//   gel<id>{vm,wm}<t>.dat       interior elements, setw(9) << double
//                               (default stream precision) (gel.cu:452-541)
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>

#include "glsm_internal.hpp"

using std::ofstream;
using std::setw;
using std::string;
using std::to_string;

namespace glsm {
namespace {

string join(const string& dir, const string& name) {
  if (dir.empty() || dir == ".") return name;
  return dir.back() == '/' ? dir + name : dir + "/" + name;
}

ofstream open_out(const string& path) {
  ofstream o(path);
  if (!o) throw std::runtime_error("cannot write " + path + " (missing directory?)");
  return o;
}

template <class V>
void row3(ofstream& o, const V& a) {
  o << setw(9) << to_string(a.x) << "      " << setw(9) << to_string(a.y) << "      " << setw(9) << to_string(a.z)
    << "\n";
}

}  // namespace

void write_fluid_dat(const string& dir, int time, int Nx, int Ny, int Nz, const float3* u, const float* c) {
  ofstream fV = open_out(join(dir, "Velb" + to_string(time) + ".dat"));
  ofstream fC = open_out(join(dir, "Conc" + to_string(time) + ".dat"));
  for (int z = 0; z < Nz; z++)
    for (int y = 0; y < Ny; y++)
      for (int x = 0; x < Nx; x++) {
        const size_t gi = (size_t)x + (size_t)y * Nx + (size_t)z * Nx * Ny;
        row3(fV, u[gi]);
        fC << setw(9) << to_string(c[gi]) << "\n";
      }
}

void write_gel_dat(const string& dir, int gel_id, int time, int LX, int LY, int LZ, const double3* rn,
                   const double3* Fn, const double3* Veln, const double3* rm, const double* um, const double* vm,
                   const double* wm) {
  const string id = to_string(gel_id), t = to_string(time);
  ofstream frn = open_out(join(dir, "gel" + id + "rn" + t + ".dat"));
  ofstream frm = open_out(join(dir, "gel" + id + "rm" + t + ".dat"));
  ofstream fum = open_out(join(dir, "gel" + id + "um" + t + ".dat"));
  ofstream fvm = open_out(join(dir, "gel" + id + "vm" + t + ".dat"));
  ofstream fwm = open_out(join(dir, "gel" + id + "wm" + t + ".dat"));
  ofstream fFn = open_out(join(dir, "gel" + id + "Fn" + t + ".dat"));
  ofstream fVeln = open_out(join(dir, "gel" + id + "Veln" + t + ".dat"));
  auto node = [&](int x, int y, int z) { return (size_t)x + (size_t)y * (LX + 2) + (size_t)z * (LX + 2) * (LY + 2); };
  auto elem = [&](int x, int y, int z) { return (size_t)x + (size_t)y * (LX + 1) + (size_t)z * (LX + 1) * (LY + 1); };
  for (int z = 1; z < LZ + 1; z++)
    for (int y = 1; y < LY + 1; y++)
      for (int x = 1; x < LX + 1; x++) {
        const size_t gi = node(x, y, z);
        row3(frn, rn[gi]);
        row3(fFn, Fn[gi]);
        row3(fVeln, Veln[gi]);
      }
  for (int z = 1; z < LZ; z++)
    for (int y = 1; y < LY; y++)
      for (int x = 1; x < LX; x++) {
        const size_t gi = elem(x, y, z);
        row3(frm, rm[gi]);
        fum << setw(9) << to_string(um[gi]) << "\n";
        fvm << setw(9) << vm[gi] << "\n";
        fwm << setw(9) << wm[gi] << "\n";
      }
}

}  // namespace glsm
