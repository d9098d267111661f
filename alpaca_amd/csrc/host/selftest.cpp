// Command-line driver of the host adapters (used by the GPU tests):
//
//   alpaca_selftest run <in.bin> <out.bin>
//       GpuUniformSolver: in = header + domain conservatives; out = steps,
//       final time, final domain state.
//   alpaca_selftest fluxes <in.bin> <out.bin>
//       GpuConvectiveTermSolver::UpdateConvectiveFluxes on one haloed block;
//       out = the three accumulated face-flux arrays.
//
// Binary header (little endian): int32 dim, nb[3], bc[3], flux, precision,
// max_steps; float64 gamma, background_pressure, cell_size, cfl, end_time.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "alpaca_host.h"

namespace {

struct Header {
  int32_t dim, nb[3], bc[3], flux, precision, max_steps;
  double gamma, b, cell_size, cfl, end_time;
};

std::vector<double> ReadRest(std::ifstream &in) {
  std::vector<double> v;
  double x;
  while (in.read(reinterpret_cast<char *>(&x), sizeof(x))) v.push_back(x);
  return v;
}

template <int DIM> int Run(Header const &h, std::vector<double> const &domain, char const *out_path) {
  alpaca_case_desc const d = alpaca_host::MakeDesc(DIM, h.nb, h.bc, h.flux, h.precision, h.gamma,
                                                   h.b, h.cell_size);
  alpaca_host::GpuUniformSolver<DIM> solver(d, h.cfl, h.end_time);
  solver.Initialize(domain);
  int const steps = solver.ComputeLoop(h.max_steps);
  alpaca_case_state const s = solver.Blocks().ReadState();
  std::vector<double> const st = solver.State();
  std::ofstream out(out_path, std::ios::binary);
  int32_t const n = steps;
  out.write(reinterpret_cast<char const *>(&n), sizeof(n));
  out.write(reinterpret_cast<char const *>(&s.time), sizeof(s.time));
  out.write(reinterpret_cast<char const *>(st.data()), st.size() * sizeof(double));
  return 0;
}

template <int DIM> int Fluxes(Header const &h, std::vector<double> const &block, char const *out_path) {
  using G = alpaca_host::Geometry<DIM>;
  if (block.size() != G::kBlockDoubles) {
    std::cerr << "block size mismatch\n";
    return 2;
  }
  alpaca_host::GpuConvectiveTermSolver<DIM> solver(h.flux, h.gamma, h.b, h.precision);
  auto *avg = new typename G::Conservatives[1];
  std::memcpy(avg, block.data(), sizeof(typename G::Conservatives));
  auto *fx = new typename G::FaceFluxes[3];
  std::memset(fx, 0, 3 * sizeof(typename G::FaceFluxes));
  auto *vf = new typename G::VolumeForces[1];
  std::memset(vf, 0, sizeof(typename G::VolumeForces));
  solver.UpdateConvectiveFluxes(avg[0], h.cell_size, fx[0], fx[1], fx[2], vf[0]);
  std::ofstream out(out_path, std::ios::binary);
  out.write(reinterpret_cast<char const *>(fx), 3 * sizeof(typename G::FaceFluxes));
  delete[] avg;
  delete[] fx;
  delete[] vf;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) {
    std::cerr << "usage: alpaca_selftest run|fluxes <in.bin> <out.bin>\n";
    return 2;
  }
  std::ifstream in(argv[2], std::ios::binary);
  Header h{};
  if (!in.read(reinterpret_cast<char *>(&h), sizeof(h))) {
    std::cerr << "cannot read header\n";
    return 2;
  }
  std::vector<double> const data = ReadRest(in);
  try {
    bool const run = std::strcmp(argv[1], "run") == 0;
    if (h.dim == 2) return run ? Run<2>(h, data, argv[3]) : Fluxes<2>(h, data, argv[3]);
    if (h.dim == 3) return run ? Run<3>(h, data, argv[3]) : Fluxes<3>(h, data, argv[3]);
    std::cerr << "bad dim\n";
    return 2;
  } catch (alpaca_host::AlpacaGpuError const &e) {
    std::cerr << "alpaca error: " << e.what() << "\n";
    return 3;
  }
}
