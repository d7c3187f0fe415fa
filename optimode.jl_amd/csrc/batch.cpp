// omx_solve_k_batch: independent solve_k problems of uneven grid, omega and mode count.
// Problems are bucketed by grid; each bucket runs on one handle and one stream, in launch
// groups of as many problems as fit the device memory: every band of every problem in a
// group is an instance of the same LOBPCG / Newton / refinement launches
// (Solver::solve_k_multi), so the per-iteration launches and the small dense steps are
// shared by all of them instead of repeated per problem.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "solver.hpp"

namespace omx {

namespace {

// device bytes of a launch group of `ninst` instances of block size M (both LOBPCG batches,
// the refinement and correction-solve vectors)
size_t group_bytes(long N, long ld, long ninst, long M) {
  const size_t lob = size_t(ninst) * M * (11 * size_t(ld) + 3 * size_t(N));
  return (16 + 8) * lob + size_t(ninst) * (16 * 16 * size_t(ld) + 16 * 3 * size_t(N));
}

}  // namespace

void solve_k_batch(int device, int nprob, omx_problem* probs, int max_group, int fp64_only) {
  if (nprob < 0 || (nprob > 0 && !probs)) throw Error(OMX_ERR_INVALID_ARG, "bad problem list");
  if (nprob == 0) return;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    throw Error(OMX_ERR_NO_DEVICE, "no such HIP device");
  for (int j = 0; j < nprob; ++j) {
    probs[j].status = OMX_ERR_INTERNAL;
    std::memset(&probs[j].info, 0, sizeof(probs[j].info));
  }
  OMX_HIP_CHECK(hipSetDevice(device));
  size_t free_b = 0, total_b = 0;
  OMX_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const size_t budget = size_t(0.5 * double(free_b));
  // buckets of one grid, problems by decreasing nev (similar block sizes share a group)
  std::map<std::tuple<int, int, double, double>, std::vector<int>> buckets;
  for (int j = 0; j < nprob; ++j) {
    const omx_problem& p = probs[j];
    if (!p.epsinv || !p.kmags_out || p.nx < 2 || p.ny < 2) {
      probs[j].status = (p.nx < 2 || p.ny < 2) ? OMX_ERR_BAD_SHAPE : OMX_ERR_INVALID_ARG;
      continue;
    }
    buckets[std::make_tuple(p.nx, p.ny, p.Dx, p.Dy)].push_back(j);
  }
  for (auto& kv : buckets) {
    std::vector<int>& ids = kv.second;
    std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return probs[a].nev > probs[b].nev; });
    const omx_problem& p0 = probs[ids[0]];
    std::unique_ptr<Solver> h;
    try {
      h = std::make_unique<Solver>(device, p0.nx, p0.ny, p0.Dx, p0.Dy, p0.epsinv);
      h->opt.fp64_only = fp64_only ? 1 : 0;
    } catch (const Error& e) {
      for (int j : ids) probs[j].status = e.code();
      continue;
    } catch (...) {
      for (int j : ids) probs[j].status = OMX_ERR_INTERNAL;
      continue;
    }
    size_t at = 0;
    long last_ninst = -1, last_M = -1;
    while (at < ids.size()) {
      // the largest group from `at` that fits the memory budget (at least one problem)
      long ninst = 0, M = 0;
      size_t end = at;
      while (end < ids.size()) {
        const omx_problem& q = probs[ids[end]];
        const long n2 = ninst + q.nev, M2 = std::max<long>(M, q.nev);
        if (end > at && (group_bytes(h->N, h->ld, n2, M2) > budget || (max_group > 0 && long(end - at) >= max_group)))
          break;
        ninst = n2;
        M = M2;
        ++end;
      }
      if (ninst != last_ninst || M != last_M) h->workspaces.clear();  // shapes change: free the old batches
      last_ninst = ninst;
      last_M = M;
      std::vector<MultiProblem> group;
      for (size_t g = at; g < end; ++g) {
        const omx_problem& q = probs[ids[g]];
        group.push_back(MultiProblem{q.epsinv, q.omega, q.nev, q.k_tol, q.eig_tol, q.max_eigsolves, q.maxiter,
                                     q.kmags_out, q.evecs_out, omx_info{}, OMX_ERR_INTERNAL});
      }
      try {
        h->solve_k_multi(group);
        for (size_t g = at; g < end; ++g) {
          probs[ids[g]].info = group[g - at].info;
          probs[ids[g]].status = group[g - at].status;
        }
      } catch (const Error& e) {
        for (size_t g = at; g < end; ++g) probs[ids[g]].status = e.code();
      } catch (...) {
        for (size_t g = at; g < end; ++g) probs[ids[g]].status = OMX_ERR_INTERNAL;
      }
      at = end;
    }
  }
  for (int j = 0; j < nprob; ++j)
    if (probs[j].status != OMX_OK)
      throw Error(probs[j].status, "omx_solve_k_batch: problem " + std::to_string(j) + " failed");
}

}  // namespace omx
