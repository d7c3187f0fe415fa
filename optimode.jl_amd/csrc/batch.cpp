// omx_solve_k_batch: independent solve_k problems of uneven grid, omega and mode count run
// concurrently from a shared queue.  Each host worker owns its handles (one per grid shape,
// each with its own HIP stream), so that the latency-bound steps of one solve overlap the
// kernels of others on the GPU.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <thread>
#include <tuple>
#include <vector>

#include "solver.hpp"

namespace omx {

void solve_k_batch(int device, int nprob, omx_problem* probs, int nstreams, int fp64_only) {
  if (nprob < 0 || (nprob > 0 && !probs)) throw Error(OMX_ERR_INVALID_ARG, "bad problem list");
  if (nprob == 0) return;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    throw Error(OMX_ERR_NO_DEVICE, "no such HIP device");
  for (int j = 0; j < nprob; ++j) probs[j].status = OMX_ERR_INTERNAL;
  // largest first: N * nev as the cost estimate
  std::vector<int> order(nprob);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return double(probs[a].nx) * probs[a].ny * probs[a].nev > double(probs[b].nx) * probs[b].ny * probs[b].nev;
  });
  const int nw = std::max(1, std::min({nstreams, nprob, 8}));
  OMX_HIP_CHECK(hipSetDevice(device));
  size_t free_b = 0, total_b = 0;
  OMX_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const size_t budget = size_t(0.6 * double(free_b)) / size_t(nw);  // LOBPCG workspaces per worker
  std::atomic<int> next{0};
  auto worker = [&]() {
    if (hipSetDevice(device) != hipSuccess) return;
    std::map<std::tuple<int, int, double, double>, std::unique_ptr<Solver>> handles;
    for (;;) {
      const int j = next.fetch_add(1);
      if (j >= nprob) break;
      omx_problem& p = probs[order[j]];
      std::memset(&p.info, 0, sizeof(p.info));
      try {
        if (!p.epsinv || !p.kmags_out) throw Error(OMX_ERR_INVALID_ARG, "NULL epsinv or kmags_out");
        auto key = std::make_tuple(p.nx, p.ny, p.Dx, p.Dy);
        auto& h = handles[key];
        if (!h)
          h = std::make_unique<Solver>(device, p.nx, p.ny, p.Dx, p.Dy, p.epsinv);
        else
          h->set_epsinv(p.epsinv);
        h->opt.fp64_only = fp64_only ? 1 : 0;
        // keep this worker's workspaces within its share of the device memory
        size_t held = 0;
        for (auto& kv : handles) held += kv.second ? kv.second->workspace_bytes() : 0;
        if (held > budget)
          for (auto& kv : handles)
            if (kv.second && kv.second.get() != h.get()) kv.second->workspaces.clear();
        if (h->workspace_bytes() > budget) h->workspaces.clear();
        h->solve_k(p.omega, p.nev, p.k_tol, p.eig_tol, p.max_eigsolves, p.maxiter, nullptr, nullptr, 0, p.kmags_out,
                   p.evecs_out, &p.info);
        p.status = OMX_OK;
      } catch (const Error& e) {
        p.status = e.code();
      } catch (...) {
        p.status = OMX_ERR_INTERNAL;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int w = 0; w < nw; ++w) pool.emplace_back(worker);
  for (auto& t : pool) t.join();
  for (int j = 0; j < nprob; ++j)
    if (probs[j].status != OMX_OK)
      throw Error(probs[j].status, "omx_solve_k_batch: problem " + std::to_string(j) + " failed");
}

}  // namespace omx
