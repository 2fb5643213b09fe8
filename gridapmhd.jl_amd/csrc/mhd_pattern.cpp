// Sparsity pattern of the H1-HDiv Jacobian (host side, built once per context).
//
// Gridap's SparseMatrixAssembler keeps, for every cell, every (row, col) of
// every field block the weak form touches (src/weakforms.jl:311: uu, up, uj,
// pu, jj, jphi, ju, phij); Dirichlet rows and columns are dropped.  The
// pattern is structurally symmetric, so one CSR (sorted columns) serves as
// CSR of A and, read column-wise, as Gridap's CSC of A.  The skeleton
// stabilisation adds, per interior face, the same-component velocity couplings
// of the face's two cells (the jump(grad u) : jump(grad v) face integrand).
#include <algorithm>
#include <atomic>
#include <thread>

#include "mhd_internal.h"

namespace mhd {
namespace {

inline int field_of(int l) { return l < OFF_P ? 0 : (l < OFF_J ? 1 : (l < OFF_PHI ? 2 : 3)); }

constexpr bool kTouched[4][4] = {
    {true, true, true, false},
    {true, false, false, false},
    {true, false, true, true},
    {false, false, true, false},
};

template <class F>
void parallel_chunks(int64_t n, int64_t chunk, F &&f) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::atomic<int64_t> next{0};
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (;;) {
        int64_t b = next.fetch_add(chunk);
        if (b >= n) break;
        f(t, b, std::min(n, b + chunk));
      }
    });
  for (auto &x : th) x.join();
}

}  // namespace

void build_pattern(int32_t ncells, const int32_t *cell_dofs, int64_t n, std::vector<int64_t> &rowptr,
                   std::vector<int32_t> &colind, int32_t nskel, const int32_t *skel_dofs) {
  // dof -> (cell, local) incidence
  std::vector<int64_t> inc_ptr(n + 1, 0);
  for (int64_t c = 0; c < ncells; ++c)
    for (int l = 0; l < NDOF; ++l) {
      int32_t g = cell_dofs[c * NDOF + l];
      if (g >= 0) inc_ptr[g + 1]++;
    }
  for (int64_t i = 0; i < n; ++i) inc_ptr[i + 1] += inc_ptr[i];
  std::vector<int32_t> inc_cell(inc_ptr[n]);
  std::vector<uint8_t> inc_loc(inc_ptr[n]);
  {
    std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int32_t c = 0; c < ncells; ++c)
      for (int l = 0; l < NDOF; ++l) {
        int32_t g = cell_dofs[(int64_t)c * NDOF + l];
        if (g >= 0) {
          int64_t p = fill[g]++;
          inc_cell[p] = c;
          inc_loc[p] = (uint8_t)l;
        }
      }
  }
  // dof -> (skeleton face, component slot) incidence
  std::vector<int64_t> sinc_ptr(skel_dofs ? n + 1 : 1, 0);
  std::vector<int32_t> sinc;
  if (skel_dofs) {
    for (int64_t k = 0; k < (int64_t)nskel * SKEL_DOFS; ++k)
      if (skel_dofs[k] >= 0) sinc_ptr[skel_dofs[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) sinc_ptr[i + 1] += sinc_ptr[i];
    sinc.resize(sinc_ptr[n]);
    std::vector<int64_t> fill(sinc_ptr.begin(), sinc_ptr.end() - 1);
    for (int64_t k = 0; k < (int64_t)nskel * SKEL_DOFS; ++k)
      if (skel_dofs[k] >= 0) sinc[fill[skel_dofs[k]]++] = (int32_t)(k / SKEL_SLOTS);  // face * 3 + component
  }
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::vector<int32_t>> marks(nt);
  std::vector<std::vector<int32_t>> bufs(nt);
  auto row_cols = [&](unsigned t, int64_t i, std::vector<int32_t> &out) {
    auto &mark = marks[t];
    if (mark.empty()) mark.assign(n, -1);
    out.clear();
    for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; ++p) {
      const int32_t c = inc_cell[p];
      const int ft = field_of(inc_loc[p]);
      const int32_t *cd = cell_dofs + (int64_t)c * NDOF;
      for (int l = 0; l < NDOF; ++l) {
        if (!kTouched[ft][field_of(l)]) continue;
        const int32_t g = cd[l];
        if (g < 0 || mark[g] == (int32_t)i) continue;
        mark[g] = (int32_t)i;
        out.push_back(g);
      }
    }
    if (skel_dofs)
      for (int64_t p = sinc_ptr[i]; p < sinc_ptr[i + 1]; ++p) {
        const int32_t *sd = skel_dofs + (int64_t)sinc[p] * SKEL_SLOTS;
        for (int v = 0; v < SKEL_SLOTS; ++v) {
          const int32_t g = sd[v];
          if (g < 0 || mark[g] == (int32_t)i) continue;
          mark[g] = (int32_t)i;
          out.push_back(g);
        }
      }
    std::sort(out.begin(), out.end());
  };
  rowptr.assign(n + 1, 0);
  parallel_chunks(n, 4096, [&](unsigned t, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      row_cols(t, i, bufs[t]);
      rowptr[i + 1] = (int64_t)bufs[t].size();
    }
  });
  for (int64_t i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
  colind.resize(rowptr[n]);
  for (auto &m : marks) std::fill(m.begin(), m.end(), -1);
  parallel_chunks(n, 4096, [&](unsigned t, int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      row_cols(t, i, bufs[t]);
      std::copy(bufs[t].begin(), bufs[t].end(), colind.begin() + rowptr[i]);
    }
  });
}

}  // namespace mhd
