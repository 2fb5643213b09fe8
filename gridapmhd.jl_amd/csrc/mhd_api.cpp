// C-ABI entry points of libmhd (declared in include/mhd.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>

#include "mhd_internal.h"

namespace {

int copy_phys(mhd_ctx *c, const mhd_params *p) {
  MhdPhys &ph = c->phys;
  ph.alpha = p->alpha;
  ph.beta = p->beta;
  ph.gamma = p->gamma;
  ph.sigma = p->sigma;
  for (int i = 0; i < 3; ++i) {
    ph.B[i] = p->B[i];
    ph.f0[i] = p->f0[i];
    ph.g0[i] = p->g0[i];
  }
  for (int i = 0; i < 9; ++i) {
    ph.fx[i] = p->fx[i];
    ph.gx[i] = p->gx[i];
  }
  ph.zeta_u = p->zeta_u;
  ph.zeta_j = p->zeta_j;
  if (p->convection < 0 || p->convection > 2) {
    c->err = "convection must be 0 (none), 1 (picard) or 2 (newton)";
    return -1;
  }
  ph.convection = p->convection;
  return 0;
}

template <class T>
int dev_upload(mhd_ctx *c, T **dst, const T *src, size_t count) {
  size_t bytes = sizeof(T) * std::max<size_t>(count, 1);
  MHD_HIP(hipMalloc(reinterpret_cast<void **>(dst), bytes));
  if (src && count) MHD_HIP(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
  else MHD_HIP(hipMemset(*dst, 0, bytes));
  return 0;
}


}  // namespace

void mhd_solver_free(mhd_ctx *c);  // mhd_solver.hip

extern "C" {

const char *mhd_last_error(const mhd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mhd_create_stabilized(mhd_ctx **out, const mhd_mesh *mesh, const mhd_spaces *spaces, const mhd_params *params,
                          const mhd_skeleton *skel, int32_t device) {
  if (!out || !mesh || !spaces || !params) return -1;
  *out = nullptr;
  mhd_ctx *c = new (std::nothrow) mhd_ctx();
  if (!c) return -4;
  *out = c;
  c->device = device;
  MHD_HIP(hipSetDevice(device));
  MHD_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  if (mesh->ncells <= 0 || !mesh->cell_xyz || !spaces->cell_dofs) {
    c->err = "mesh/spaces: empty or missing arrays";
    return -1;
  }
  if (mesh->nfaces_tw < 0 || (mesh->nfaces_tw > 0 && (!mesh->tw_face_cell || !mesh->tw_face_lface || !mesh->tw_tau ||
                                                      !mesh->tw_cw || !mesh->tw_jw))) {
    c->err = "mesh: thin-wall face arrays missing";
    return -1;
  }
  int rc = copy_phys(c, params);
  if (rc) return rc;
  c->ncells = mesh->ncells;
  c->nfaces = mesh->nfaces_tw;
  c->ndir = spaces->ndir;
  int64_t n = 0;
  for (int f = 0; f < 4; ++f) {
    if (spaces->nfree[f] < 0) {
      c->err = "spaces: negative field size";
      return -1;
    }
    c->nfree[f] = spaces->nfree[f];
    n += spaces->nfree[f];
  }
  c->n = n;
  // validate the dof table on the host before anything reaches a kernel
  const int64_t ntab = (int64_t)c->ncells * mhd::NDOF;
  for (int64_t k = 0; k < ntab; ++k) {
    const int32_t g = spaces->cell_dofs[k];
    if (g >= n || (g < 0 && -(int64_t)g - 1 >= std::max<int64_t>(c->ndir, 0))) {
      c->err = "spaces: cell_dofs entry out of range at " + std::to_string(k);
      return -1;
    }
  }
  for (int f = 0; f < c->nfaces; ++f) {
    const int32_t cell = mesh->tw_face_cell[f];
    const int lf = mesh->tw_face_lface[f];
    if (cell < 0 || cell >= c->ncells || lf < 0 || lf > 5) {
      c->err = "mesh: thin-wall face " + std::to_string(f) + " out of range";
      return -1;
    }
  }
  if (c->ndir > 0 && !spaces->dir_vals) {
    c->err = "spaces: dir_vals missing";
    return -1;
  }
  // skeleton faces: slot dofs of the two cells' velocity nodes, checked on the host
  std::vector<int32_t> sdofs;
  std::vector<uint8_t> mslot;
  const int32_t nskel = skel ? skel->nfaces : 0;
  if (nskel < 0 || (nskel > 0 && (!skel->face_cell || !skel->face_lface || !skel->face_perm || !skel->mu))) {
    c->err = "skeleton: face arrays missing";
    return -1;
  }
  if ((int64_t)nskel * mhd::SKEL_DOFS >= ((int64_t)1 << 31)) {
    c->err = "skeleton: too many faces for the int32 entry index";
    return -1;
  }
  if (nskel > 0) {
    sdofs.assign((size_t)nskel * mhd::SKEL_DOFS, mhd::SKEL_NONE);
    mslot.assign((size_t)nskel * 27, 0);
    for (int32_t f = 0; f < nskel; ++f) {
      const std::string where = "skeleton face " + std::to_string(f) + ": ";
      const int32_t c0 = skel->face_cell[2 * f], c1 = skel->face_cell[2 * f + 1];
      const int lf0 = skel->face_lface[2 * f], lf1 = skel->face_lface[2 * f + 1];
      if (c0 < 0 || c0 >= c->ncells || c1 < 0 || c1 >= c->ncells || lf0 < 0 || lf0 > 5 || lf1 < 0 || lf1 > 5) {
        c->err = where + "cell or local face out of range";
        return -1;
      }
      if (c0 == c1) {
        c->err = where + "a cell glued to itself (periodic direction one cell thick) is not supported";
        return -1;
      }
      if (!std::isfinite(skel->mu[f])) {
        c->err = where + "mu is not finite";
        return -1;
      }
      int seen = 0;
      for (int q = 0; q < mhd::NQF; ++q) {
        const int p = skel->face_perm[(int64_t)9 * f + q];
        if (p < 0 || p >= mhd::NQF || (seen >> p & 1)) {
          c->err = where + "face_perm is not a permutation of 0..8";
          return -1;
        }
        seen |= 1 << p;
      }
      const int32_t *d0 = spaces->cell_dofs + (int64_t)c0 * mhd::NDOF;
      const int32_t *d1 = spaces->cell_dofs + (int64_t)c1 * mhd::NDOF;
      int32_t *sd = sdofs.data() + (int64_t)f * mhd::SKEL_DOFS;
      for (int a = 0; a < 27; ++a) {
        for (int b = 0; b < a; ++b)
          if (d0[b] == d0[a]) {
            c->err = where + "a cell holds a velocity node twice";
            return -1;
          }
        for (int comp = 0; comp < 3; ++comp) sd[comp * mhd::SKEL_SLOTS + a] = d0[comp * 27 + a];
      }
      // node a = i + 3j + 9k lies on local face lf (z-, z+, y-, y+, x-, x+) when
      // its coordinate along the face axis is 0 (lower side) or 2 (upper side)
      auto on_face = [](int a, int lf) {
        const int axis = 2 - lf / 2, coord = axis == 0 ? a % 3 : (axis == 1 ? (a / 3) % 3 : a / 9);
        return coord == ((lf & 1) ? 2 : 0);
      };
      int matched = 0;
      for (int a1 = 0; a1 < 27; ++a1) {
        int slot = -1;
        for (int a0 = 0; a0 < 27 && slot < 0; ++a0)
          if (d0[a0] == d1[a1]) slot = a0;
        if (slot >= 0) {
          if (on_face(a1, lf1) && on_face(slot, lf0)) ++matched;
        } else {
          slot = 27 + a1;
          for (int comp = 0; comp < 3; ++comp) sd[comp * mhd::SKEL_SLOTS + slot] = d1[comp * 27 + a1];
        }
        mslot[(size_t)f * 27 + a1] = (uint8_t)slot;
      }
      if (matched != mhd::NQF) {
        c->err = where + "the two cells do not share the nodes of the given local faces";
        return -1;
      }
    }
  }
  c->block_style = spaces->block_style;
  c->h_dofs.assign(spaces->cell_dofs, spaces->cell_dofs + ntab);
  c->h_xyz.assign(mesh->cell_xyz, mesh->cell_xyz + (size_t)c->ncells * 24);
  if (mesh->cell_verts) {
    c->nverts = mesh->nverts;
    c->h_verts.assign(mesh->cell_verts, mesh->cell_verts + (size_t)c->ncells * 8);
    for (int32_t v : c->h_verts)
      if (v < 0 || v >= c->nverts) {
        c->err = "mesh: cell_verts entry out of range";
        return -1;
      }
  }
  if ((rc = dev_upload(c, &c->d_xyz, mesh->cell_xyz, (size_t)c->ncells * 24))) return rc;
  if ((rc = dev_upload(c, &c->d_dofs, spaces->cell_dofs, (size_t)ntab))) return rc;
  if (spaces->rt_sign) {
    if ((rc = dev_upload(c, &c->d_sign, spaces->rt_sign, (size_t)c->ncells * 6))) return rc;
  }
  if ((rc = dev_upload(c, &c->d_dir, spaces->dir_vals, (size_t)std::max(c->ndir, 0)))) return rc;
  if ((rc = dev_upload(c, &c->d_fcell, mesh->tw_face_cell, (size_t)c->nfaces))) return rc;
  if ((rc = dev_upload(c, &c->d_flf, mesh->tw_face_lface, (size_t)c->nfaces))) return rc;
  if ((rc = dev_upload(c, &c->d_ftau, mesh->tw_tau, (size_t)c->nfaces))) return rc;
  if ((rc = dev_upload(c, &c->d_fcw, mesh->tw_cw, (size_t)c->nfaces))) return rc;
  if ((rc = dev_upload(c, &c->d_fjw, mesh->tw_jw, (size_t)c->nfaces))) return rc;
  {
    mhd::RefTables *T = new mhd::RefTables;
    mhd::FaceTables *F = new mhd::FaceTables;
    mhd::build_ref_tables(T);
    mhd::build_face_tables(F);
    rc = dev_upload(c, &c->d_ref, T, 1);
    if (!rc) rc = dev_upload(c, &c->d_fref, F, 1);
    delete T;
    delete F;
    if (rc) return rc;
  }
  // sparsity pattern (host) and cell -> nonzero rank map (device)
  mhd::build_pattern(c->ncells, spaces->cell_dofs, n, c->h_rowptr, c->h_colind, nskel,
                     nskel > 0 ? sdofs.data() : nullptr);
  c->nnz = c->h_rowptr[n];
  if ((rc = dev_upload(c, &c->d_rowptr, c->h_rowptr.data(), (size_t)n + 1))) return rc;
  if ((rc = dev_upload(c, &c->d_colind, c->h_colind.data(), (size_t)c->nnz))) return rc;
  MHD_HIP(hipMalloc(&c->d_rank, sizeof(uint16_t) * (size_t)c->ncells * mhd::NENT));
  if ((rc = mhd_build_rank_map(c))) return rc;
  // greedy colouring of the cell / shared-dof graph
  {
    std::vector<int64_t> ptr(n + 1, 0);
    for (int64_t k = 0; k < ntab; ++k)
      if (spaces->cell_dofs[k] >= 0) ptr[spaces->cell_dofs[k] + 1]++;
    for (int64_t i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
    std::vector<int32_t> cells_of(ptr[n]);
    {
      std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
      for (int32_t cc = 0; cc < c->ncells; ++cc)
        for (int l = 0; l < mhd::NDOF; ++l) {
          const int32_t g = spaces->cell_dofs[(int64_t)cc * mhd::NDOF + l];
          if (g >= 0) cells_of[fill[g]++] = cc;
        }
    }
    std::vector<int32_t> color(c->ncells, -1);
    std::vector<int32_t> seen(64, -1);
    int32_t ncol = 0;
    for (int32_t cc = 0; cc < c->ncells; ++cc) {
      for (int l = 0; l < mhd::NDOF; ++l) {
        const int32_t g = spaces->cell_dofs[(int64_t)cc * mhd::NDOF + l];
        if (g < 0) continue;
        for (int64_t q = ptr[g]; q < ptr[g + 1]; ++q) {
          const int32_t o = color[cells_of[q]];
          if (o >= 0) {
            if (o >= (int32_t)seen.size()) seen.resize(o + 1, -1);
            seen[o] = cc;
          }
        }
      }
      int32_t k = 0;
      while (k < (int32_t)seen.size() && seen[k] == cc) ++k;
      color[cc] = k;
      ncol = std::max(ncol, k + 1);
    }
    c->color_ptr.assign(ncol + 1, 0);
    for (int32_t cc = 0; cc < c->ncells; ++cc) c->color_ptr[color[cc] + 1]++;
    for (int32_t k = 0; k < ncol; ++k) c->color_ptr[k + 1] += c->color_ptr[k];
    std::vector<int32_t> sorted(c->ncells), fill(c->color_ptr.begin(), c->color_ptr.end() - 1);
    for (int32_t cc = 0; cc < c->ncells; ++cc) sorted[fill[color[cc]]++] = cc;
    if ((rc = dev_upload(c, &c->d_color_cells, sorted.data(), sorted.size()))) return rc;
    // row assembly lists: cells_of is already in increasing cell order per dof
    if ((int64_t)c->ncells * mhd::NDOF >= ((int64_t)1 << 31)) {
      c->err = "too many cells for the row-assembly index (int32)";
      return -1;
    }
    std::vector<int32_t> crow(n + 1), centry(ptr[n]);
    for (int64_t i = 0; i <= n; ++i) crow[i] = (int32_t)ptr[i];
    {
      std::vector<int64_t> fill2(ptr.begin(), ptr.end() - 1);
      for (int32_t cc = 0; cc < c->ncells; ++cc)
        for (int l = 0; l < mhd::NDOF; ++l) {
          const int32_t g = spaces->cell_dofs[(int64_t)cc * mhd::NDOF + l];
          if (g >= 0) centry[fill2[g]++] = cc * mhd::NDOF + l;
        }
    }
    if ((rc = dev_upload(c, &c->d_crow, crow.data(), crow.size()))) return rc;
    if (centry.empty()) centry.push_back(0);
    if ((rc = dev_upload(c, &c->d_centry, centry.data(), centry.size()))) return rc;
    // a dof listed twice in one cell (periodic direction one cell thick)
    for (int32_t cc = 0; cc < c->ncells && !c->dup_dofs; ++cc) {
      int32_t v[mhd::NDOF];
      int m = 0;
      for (int l = 0; l < mhd::NDOF; ++l) {
        const int32_t g = spaces->cell_dofs[(int64_t)cc * mhd::NDOF + l];
        if (g >= 0) v[m++] = g;
      }
      std::sort(v, v + m);
      for (int l = 1; l < m; ++l)
        if (v[l] == v[l - 1]) c->dup_dofs = true;
    }
  }
  for (int64_t i = 0; i < n; ++i)
    c->max_row = std::max<int32_t>(c->max_row, (int32_t)(c->h_rowptr[i + 1] - c->h_rowptr[i]));
  MHD_HIP(hipMalloc(&c->d_val, sizeof(double) * std::max<int64_t>(c->nnz, 1)));
  MHD_HIP(hipMalloc(&c->d_x, sizeof(double) * std::max<int64_t>(n, 1)));
  MHD_HIP(hipMalloc(&c->d_r, sizeof(double) * std::max<int64_t>(n, 1)));
  if (nskel > 0) {
    // skeleton entries of every row, in increasing face order
    std::vector<int32_t> srow(n + 1, 0);
    for (int32_t g : sdofs)
      if (g >= 0) srow[g + 1]++;
    for (int64_t i = 0; i < n; ++i) srow[i + 1] += srow[i];
    std::vector<int32_t> sentry(std::max<int32_t>(srow[n], 1)), fill(srow.begin(), srow.end() - 1);
    for (int64_t k = 0; k < (int64_t)sdofs.size(); ++k)
      if (sdofs[k] >= 0) sentry[fill[sdofs[k]]++] = (int32_t)k;
    if (c->max_row >= mhd::RANK_NONE) {
      c->err = "skeleton: a CSR row of " + std::to_string(c->max_row) + " entries exceeds the 16-bit rank table";
      return -1;
    }
    c->nskel = nskel;
    c->nsentry = srow[n];
    if ((rc = dev_upload(c, &c->d_skel_dofs, sdofs.data(), sdofs.size()))) return rc;
    if ((rc = dev_upload(c, &c->d_srow, srow.data(), srow.size()))) return rc;
    if ((rc = dev_upload(c, &c->d_sentry, sentry.data(), sentry.size()))) return rc;
    int32_t *d_fc = nullptr;
    int8_t *d_lf = nullptr, *d_perm = nullptr;
    uint8_t *d_ms = nullptr;
    double *d_mu = nullptr;
    rc = dev_upload(c, &d_fc, skel->face_cell, (size_t)2 * nskel);
    if (!rc) rc = dev_upload(c, &d_lf, skel->face_lface, (size_t)2 * nskel);
    if (!rc) rc = dev_upload(c, &d_perm, skel->face_perm, (size_t)9 * nskel);
    if (!rc) rc = dev_upload(c, &d_ms, mslot.data(), mslot.size());
    if (!rc) rc = dev_upload(c, &d_mu, skel->mu, (size_t)nskel);
    if (!rc) rc = mhd_build_skeleton(c, d_fc, d_lf, d_perm, d_ms, d_mu);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) {
      c->err = "skeleton face matrices: stream synchronisation failed";
      rc = -2;
    }
    for (void *p : {(void *)d_fc, (void *)d_lf, (void *)d_perm, (void *)d_ms, (void *)d_mu})
      if (p) (void)hipFree(p);
    if (rc) return rc;
  }
  // device address of this GPU's allocation-free residual workspace (looked up
  // once per device, with the context's device current)
  return mhd_noalloc_init(c);
}

int mhd_create(mhd_ctx **out, const mhd_mesh *mesh, const mhd_spaces *spaces, const mhd_params *params,
               int32_t device) {
  return mhd_create_stabilized(out, mesh, spaces, params, nullptr, device);
}

void mhd_destroy(mhd_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  mhd_solver_free(c);
  mhd_coarse_free(c);
  void *ptrs[] = {c->d_xyz, c->d_dofs, c->d_sign, c->d_dir, c->d_fcell, c->d_flf, c->d_ftau, c->d_fcw, c->d_fjw,
                  c->d_rowptr, c->d_colind, c->d_rank, c->d_val, c->d_tperm, c->d_ref, c->d_fref, c->d_x, c->d_r,
                  c->d_color_cells, c->d_crow, c->d_centry, c->d_ebuf, c->d_rbuf, c->d_skel_dofs, c->d_skel_K,
                  c->d_srow, c->d_sentry, c->d_skel_rank, c->d_skel_res};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mhd_set_params(mhd_ctx *c, const mhd_params *p) {
  if (!c || !p) return -1;
  return copy_phys(c, p);
}

int mhd_set_dirichlet(mhd_ctx *c, const double *dir_vals) {
  if (!c) return -1;
  if (c->ndir <= 0) return 0;
  if (!dir_vals) {
    c->err = "dir_vals is NULL";
    return -1;
  }
  MHD_HIP(hipSetDevice(c->device));
  MHD_HIP(hipMemcpy(c->d_dir, dir_vals, sizeof(double) * c->ndir, hipMemcpyHostToDevice));
  return 0;
}

int mhd_set_assembly_mode(mhd_ctx *c, int32_t mode) {
  if (!c) return -1;
  if (mode < 0 || mode > 2) {
    c->err = "assembly mode must be 0 (atomic), 1 (colour classes) or 2 (element buffer + row assembly)";
    return -1;
  }
  c->assembly_mode = mode;
  return 0;
}

int mhd_sizes(mhd_ctx *c, int64_t *nfree, int64_t *nnz) {
  if (!c) return -1;
  if (nfree) *nfree = c->n;
  if (nnz) *nnz = c->nnz;
  return 0;
}

int mhd_pattern(mhd_ctx *c, int64_t *colptr, int64_t *rowval) {
  if (!c || !colptr || !rowval) return -1;
  for (int64_t i = 0; i <= c->n; ++i) colptr[i] = c->h_rowptr[i] + 1;
  for (int64_t k = 0; k < c->nnz; ++k) rowval[k] = (int64_t)c->h_colind[k] + 1;
  return 0;
}

int mhd_csr_arrays(mhd_ctx *c, const int64_t **d_rowptr, const int32_t **d_colind) {
  if (!c) return -1;
  if (d_rowptr) *d_rowptr = c->d_rowptr;
  if (d_colind) *d_colind = c->d_colind;
  return 0;
}

int mhd_residual_d(mhd_ctx *c, const double *d_x, double *d_r, void *stream) {
  if (!c || !d_x || !d_r) return -1;
  MHD_HIP(hipSetDevice(c->device));
  return mhd_launch_residual(c, d_x, d_r, stream ? (hipStream_t)stream : c->stream);
}

int mhd_residual_d_noalloc(mhd_ctx *c, const double *d_x, double *d_r, void *stream) {
  if (!c || !d_x || !d_r) return -1;
  MHD_HIP(hipSetDevice(c->device));
  return mhd_launch_residual_noalloc(c, d_x, d_r, stream ? (hipStream_t)stream : c->stream);
}

int mhd_jacobian_d(mhd_ctx *c, const double *d_x, double *d_val, void *stream) {
  if (!c || !d_x) return -1;
  MHD_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (d_val) return mhd_launch_jacobian(c, d_x, d_val, s);
  // into the context's own storage: the Jacobian mhd_precond_setup factors,
  // with its state (the coarse levels' injected states come from it)
  if (d_x != c->d_x) MHD_HIP(hipMemcpyAsync(c->d_x, d_x, sizeof(double) * c->n, hipMemcpyDeviceToDevice, s));
  int rc = mhd_launch_jacobian(c, c->d_x, c->d_val, s);
  c->jac_valid = rc == 0;
  return rc;
}

int mhd_spmv_d(mhd_ctx *c, const double *d_x, double *d_y, void *stream) {
  if (!c || !d_x || !d_y) return -1;
  MHD_HIP(hipSetDevice(c->device));
  return mhd_launch_spmv(c, c->d_val, d_x, d_y, stream ? (hipStream_t)stream : c->stream);
}

int mhd_cell_integrals(mhd_ctx *c, const double *x, double *out) {
  if (!c || !x || !out) return -1;
  MHD_HIP(hipSetDevice(c->device));
  double *d_out = nullptr;
  MHD_HIP(hipMalloc(&d_out, sizeof(double) * 6 * (size_t)c->ncells));
  MHD_HIP(hipMemcpyAsync(c->d_x, x, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  int rc = mhd_launch_cell_integrals(c, c->d_x, d_out, c->stream);
  if (!rc) {
    MHD_HIP(hipMemcpyAsync(out, d_out, sizeof(double) * 6 * (size_t)c->ncells, hipMemcpyDeviceToHost, c->stream));
    MHD_HIP(hipStreamSynchronize(c->stream));
  }
  MHD_HIP(hipFree(d_out));
  return rc;
}

int mhd_residual(mhd_ctx *c, const double *x, double *r) {
  if (!c || !x || !r) return -1;
  MHD_HIP(hipSetDevice(c->device));
  MHD_HIP(hipMemcpyAsync(c->d_x, x, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  int rc = mhd_launch_residual(c, c->d_x, c->d_r, c->stream);
  if (rc) return rc;
  MHD_HIP(hipMemcpyAsync(r, c->d_r, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  MHD_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int mhd_jacobian(mhd_ctx *c, const double *x, double *nzval) {
  if (!c || !x || !nzval) return -1;
  MHD_HIP(hipSetDevice(c->device));
  MHD_HIP(hipMemcpyAsync(c->d_x, x, sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  int rc = mhd_launch_jacobian(c, c->d_x, c->d_val, c->stream);
  c->jac_valid = rc == 0;
  if (rc) return rc;
  if (!c->d_tperm) {
    MHD_HIP(hipMalloc(&c->d_tperm, sizeof(int64_t) * std::max<int64_t>(c->nnz, 1)));
    if ((rc = mhd_build_tperm(c))) return rc;
  }
  // CSC(A) values = CSR(A) values read through the transpose permutation
  std::vector<double> csr(c->nnz);
  std::vector<int64_t> tp(c->nnz);
  MHD_HIP(hipMemcpyAsync(csr.data(), c->d_val, sizeof(double) * c->nnz, hipMemcpyDeviceToHost, c->stream));
  MHD_HIP(hipMemcpyAsync(tp.data(), c->d_tperm, sizeof(int64_t) * c->nnz, hipMemcpyDeviceToHost, c->stream));
  MHD_HIP(hipStreamSynchronize(c->stream));
  for (int64_t k = 0; k < c->nnz; ++k) nzval[k] = csr[tp[k]];
  return 0;
}

}  // extern "C"
