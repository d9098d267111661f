// Command-line driver of the host adapters (used by the GPU tests):
//
//   alpaca_selftest run <in.bin> <out.bin>
//       GpuUniformSolver: in = header + domain conservatives; out = steps,
//       final time, final domain state.
//   alpaca_selftest fluxes <in.bin> <out.bin>
//       GpuConvectiveTermSolver::UpdateConvectiveFluxes on one haloed block;
//       out = the three accumulated face-flux arrays.
//   alpaca_selftest mr <in.bin> <out.bin>
//       GpuMultiresolution on one level-0 parent and its level-1 child:
//       in = int32 dim, position, norm, pad; float64 epsilon; parent block,
//       child block, parent jump surfaces, child jump surfaces.  out = the
//       averaged parent, the child predicted over the whole block and over
//       its internal cells (every field), the detail and the decision (as a
//       float64) under GpuThresholdComputer(1, 1, epsilon), and the parent's
//       jump surfaces after AverageJumpBuffer.
//   alpaca_selftest lts <in.bin> <out.bin>
//       GpuLevelTimeIntegrator on one block: in = int32 dim, stage, n,
//       count; float64 micro[count]; avg, rhs, init blocks, jump
//       conservatives and fluxes.  out = avg, rhs, jump conservatives, jump
//       fluxes after PrepareBufferForIntegration + IntegrateNode, then the
//       rhs after IntegrateJumpHalos over the -x halo slab.
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

struct MrHeader {
  int32_t dim, position, norm, pad;
  double epsilon;
};

template <int DIM> int Mr(MrHeader const &h, std::vector<double> const &data, char const *out_path) {
  using G = alpaca_host::Geometry<DIM>;
  using M = alpaca_host::GpuMultiresolution<DIM>;
  std::size_t const nb = G::kBlockDoubles, ns = M::kSurfaceDoubles;
  if (data.size() != 2 * nb + 2 * ns) {
    std::cerr << "mr input size mismatch\n";
    return 2;
  }
  M const mr(alpaca_host::GpuThresholdComputer<DIM>(1, 1, h.epsilon), h.norm);
  std::uint64_t const child_id = (std::uint64_t(0x1400000) << 3) | std::uint64_t(h.position);
  auto *parent = new typename G::Conservatives[1];
  auto *child = new typename G::Conservatives[4];  // [1] whole block, [2] internal cells
  std::memcpy(parent, data.data(), nb * sizeof(double));
  for (int c = 0; c < 3; ++c) std::memcpy(child + c, data.data() + nb, nb * sizeof(double));
  std::vector<double> out;
  auto append = [&](void const *p, std::size_t n) {
    double const *d = static_cast<double const *>(p);
    out.insert(out.end(), d, d + n);
  };
  // ChildNeedsRemeshing first: the blocks as given.
  double detail = 0.0;
  alpaca_host::RemeshIdentifier const r = mr.ChildNeedsRemeshing(parent[0], child[0], child_id, &detail);
  typename G::Conservatives *avg = new typename G::Conservatives[1];
  std::memcpy(avg, parent, nb * sizeof(double));
  mr.Average(child[0], avg[0], child_id);
  append(avg, nb);
  for (int e = 0; e < G::NE; ++e) {
    mr.Prediction(parent[0][e], child[1][e], child_id);
    mr.Prediction(parent[0][e], child[2][e], child_id, G::HS, G::IC, G::HS, G::IC,
                  DIM == 3 ? G::HS : 0, DIM == 3 ? G::IC : 1);
  }
  append(child + 1, nb);
  append(child + 2, nb);
  append(&detail, 1);
  double const rd = double(int(r));
  append(&rd, 1);
  std::vector<double> surf(data.begin() + 2 * nb, data.begin() + 2 * nb + ns);
  mr.AverageJumpBuffer(data.data() + 2 * nb + ns, surf.data(), child_id);
  append(surf.data(), ns);
  // Jump halos: the child's halo slabs at every location of the dimension,
  // predicted from the parent in one batched call.
  {
    std::vector<alpaca_mr_link> links;
    std::vector<int> locs;
    for (int loc = 0; loc < 26; ++loc) {
      bool const z = (loc >= 4 && loc <= 13) || loc >= 18;  // a Top/Bottom part
      if (DIM == 2 && z) continue;
      links.push_back(alpaca_mr_link{1, 0, int(h.position), 1});
      locs.push_back(loc);
    }
    alpaca_host::DeviceBuffer pool(2 * nb * sizeof(double)),
        dl(links.size() * sizeof(alpaca_mr_link)), dloc(locs.size() * sizeof(int));
    alpaca_host::CheckHip(hipMemcpy(pool.get(), parent, nb * sizeof(double), hipMemcpyHostToDevice),
                          "upload");
    alpaca_host::CheckHip(hipMemcpy(pool.get<double>() + nb, data.data() + nb, nb * sizeof(double),
                                    hipMemcpyHostToDevice), "upload");
    alpaca_host::CheckHip(hipMemcpy(dl.get(), links.data(), dl.bytes(), hipMemcpyHostToDevice), "upload");
    alpaca_host::CheckHip(hipMemcpy(dloc.get(), locs.data(), dloc.bytes(), hipMemcpyHostToDevice),
                          "upload");
    mr.JumpHaloUpdateLinks(pool.get<double>(), 2, dl.get<alpaca_mr_link>(), dloc.get<int>(),
                           int(locs.size()));
    alpaca_host::CheckHip(hipMemcpy(child + 3, pool.get<double>() + nb, nb * sizeof(double),
                                    hipMemcpyDeviceToHost), "download");
    append(child + 3, nb);
  }
  std::ofstream o(out_path, std::ios::binary);
  o.write(reinterpret_cast<char const *>(out.data()), out.size() * sizeof(double));
  delete[] parent;
  delete[] child;
  delete[] avg;
  return 0;
}

struct LtsHeader {
  int32_t dim, stage, n, count;
};

template <int DIM> int Lts(LtsHeader const &h, std::vector<double> const &data, char const *out_path) {
  using G = alpaca_host::Geometry<DIM>;
  std::size_t const nb = G::kBlockDoubles;
  std::size_t const ns = std::size_t(6) * G::NE * G::IC * G::ICZ;
  if (h.count < 0 || data.size() != std::size_t(h.count) + 3 * nb + 2 * ns) {
    std::cerr << "lts input size mismatch\n";
    return 2;
  }
  alpaca_host::GpuLevelTimeIntegrator<DIM> ti;
  for (int i = 0; i < h.count; ++i) ti.AppendMicroTimestep(data[i]);
  double const *p = data.data() + h.count;
  std::size_t const bytes = sizeof(double) * (3 * nb + 2 * ns);
  alpaca_host::DeviceBuffer d(bytes), list(sizeof(int));
  alpaca_host::CheckHip(hipMemcpy(d.get(), p, bytes, hipMemcpyHostToDevice), "upload");
  int const zero = 0;
  alpaca_host::CheckHip(hipMemcpy(list.get(), &zero, sizeof(int), hipMemcpyHostToDevice), "upload");
  double *avg = d.get<double>(), *rhs = avg + nb, *init = rhs + nb, *jc = init + nb, *jf = jc + ns;
  int const *l = list.get<int>();
  ti.PrepareBufferForIntegration(avg, init, 1, l, 1, unsigned(h.stage));
  ti.IntegrateNode(avg, rhs, jc, jf, 1, l, 1, unsigned(h.stage), unsigned(h.n));
  std::vector<double> out(3 * nb + 2 * ns);
  alpaca_host::CheckHip(hipMemcpy(out.data(), d.get(), bytes, hipMemcpyDeviceToHost), "download");
  int const start[3] = {0, G::HS, DIM == 3 ? G::HS : 0}, size[3] = {G::HS, G::IC, G::ICZ};
  ti.IntegrateJumpHalos(avg, rhs, 1, l, 1, unsigned(h.stage), unsigned(h.n), start, size);
  std::vector<double> rhs2(nb);
  alpaca_host::CheckHip(hipMemcpy(rhs2.data(), rhs, nb * sizeof(double), hipMemcpyDeviceToHost),
                        "download");
  std::ofstream o(out_path, std::ios::binary);
  o.write(reinterpret_cast<char const *>(out.data()), nb * sizeof(double));           // avg
  o.write(reinterpret_cast<char const *>(out.data() + nb), nb * sizeof(double));      // rhs
  o.write(reinterpret_cast<char const *>(out.data() + 3 * nb), 2 * ns * sizeof(double));  // jc, jf
  o.write(reinterpret_cast<char const *>(rhs2.data()), nb * sizeof(double));
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) {
    std::cerr << "usage: alpaca_selftest run|fluxes|mr|lts <in.bin> <out.bin>\n";
    return 2;
  }
  std::ifstream in(argv[2], std::ios::binary);
  if (std::strcmp(argv[1], "lts") == 0) {
    LtsHeader m{};
    if (!in.read(reinterpret_cast<char *>(&m), sizeof(m))) {
      std::cerr << "cannot read header\n";
      return 2;
    }
    std::vector<double> const data = ReadRest(in);
    try {
      if (m.dim == 2) return Lts<2>(m, data, argv[3]);
      if (m.dim == 3) return Lts<3>(m, data, argv[3]);
      std::cerr << "bad dim\n";
      return 2;
    } catch (alpaca_host::AlpacaGpuError const &e) {
      std::cerr << "alpaca error: " << e.what() << "\n";
      return 3;
    }
  }
  if (std::strcmp(argv[1], "mr") == 0) {
    MrHeader m{};
    if (!in.read(reinterpret_cast<char *>(&m), sizeof(m))) {
      std::cerr << "cannot read header\n";
      return 2;
    }
    std::vector<double> const data = ReadRest(in);
    try {
      if (m.dim == 2) return Mr<2>(m, data, argv[3]);
      if (m.dim == 3) return Mr<3>(m, data, argv[3]);
      std::cerr << "bad dim\n";
      return 2;
    } catch (alpaca_host::AlpacaGpuError const &e) {
      std::cerr << "alpaca error: " << e.what() << "\n";
      return 3;
    }
  }
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
