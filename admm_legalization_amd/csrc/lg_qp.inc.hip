// QP refine of the post-process (CudaPostProcess.cu:94-316, CudaQpSolver.cu,
// PostProcessKernel.hpp:28-312, QpSolverKernels.hpp:96-475), FP64 on the GPU.
// Included at the end of lg_device.hip (same translation unit).
//
// Every cut row owns its cells and its n+1 ordering constraints, so one
// thread per cut row runs the whole ADMM iteration for that row (x-update,
// Thomas solve, constraint values, slack and dual updates); the only global
// coupling is the termination test on min(Bx - b), reduced through an
// order-preserving 64-bit key and checked by a single-thread kernel.

namespace lgdev {

struct QpArgs {
  int x_size, y_size, site_width, core_lx, N;
  int num_rows;     // cut rows
  int num_local;    // cells in cut rows
  const int16_t* pcode;
  const int32_t* stamp;     // pixel -> cell id or -1
  const int32_t* h;
  const int32_t* w;
  const int32_t* ox;
  int32_t* lx;
  const int32_t* ly;
  int32_t* row_count;       // [y_size]
  const int32_t* row_off;   // [y_size + 1]
  int32_t* cr_y;
  int32_t* cr_xb;
  int32_t* cr_xe;
  int32_t* pix_row;         // pixel -> cut row id (-1)
  int32_t* cell_row;        // [N]
  int32_t* cr_cnt;          // [num_rows + 1]
  const int32_t* cr_off;    // [num_rows + 2]
  int32_t* cr_fill;         // [num_rows + 1]
  int32_t* cr_cells;        // [N] grouped by cut row
  double* x;                // [num_local] primal (dbu)
  double* x0;               // [num_local] original lx (dbu)
  double* cp;               // [num_local] Thomas c'
  double* din;              // [num_local]
  double* dp;               // [num_local]
  double* bound;            // [num_local + num_rows]
  double* slack;
  double* dualv;
  int32_t* bl;              // DP back-pointers
  int32_t* br;
  unsigned long long* qctl; // [0] min residual key, [1] iter, [2] done, [3] iterations
  double rho;
  int max_iter, min_iter;
};

__global__ void k_qp_stamp(QpArgs q) {  // stampCellsKernel (PostProcessKernel.hpp:28-55)
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= q.N) return;
  for (int x = q.lx[c]; x < q.lx[c] + q.w[c]; x++)
    for (int y = q.ly[c]; y < q.ly[c] + q.h[c]; y++)
      atomicMax((int*)&q.stamp[(size_t)y * q.x_size + x], c);  // overlaps: highest id, as a sequential sweep
}

__device__ __forceinline__ void qp_pix(const QpArgs& q, int x, int y, int& grp, int& val, int& hh) {
  size_t p = (size_t)y * q.x_size + x;
  int code = q.pcode[p];
  val = code != 0 ? 1 : 0;
  grp = code - 2;  // only compared while valid
  int c = q.stamp[p];
  hh = c >= 0 ? q.h[c] : 0;
}

// countCutRowsKernel (PostProcessKernel.hpp:57-97)
__global__ void k_qp_count_rows(QpArgs q) {
  int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= q.y_size) return;
  int pg = -999, pv = -999, ps = -999, n = 0;
  for (int x = 0; x < q.x_size; x++) {
    int g, v, hh;
    qp_pix(q, x, y, g, v, hh);
    if (v == 0) g = -1000;  // group of an invalid pixel never decides a segment
    int single = hh <= 1 ? 1 : 0;
    bool ends = (pg != g) || (pv != v) || (hh > 1);
    if (ends && pv == 1 && ps == 1) n++;
    if (x == q.x_size - 1 && !ends && v == 1 && hh <= 1) n++;
    pg = g; pv = v; ps = single;
  }
  q.row_count[y] = n;
}

// generateCutRowsKernel (PostProcessKernel.hpp:99-161)
__global__ void k_qp_gen_rows(QpArgs q) {
  int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= q.y_size) return;
  const int base = q.row_off[y], cap = q.row_off[y + 1];
  int made = 0, pg = -999, pv = -999, ps = -999, xb = 0;
  for (int x = 0; x < q.x_size; x++) {
    int g, v, hh;
    qp_pix(q, x, y, g, v, hh);
    if (v == 0) g = -1000;
    int single = hh <= 1 ? 1 : 0;
    bool ends = (pg != g) || (pv != v) || (ps != single);
    if (ends) {
      if (pv == 1 && ps == 1 && base + made < cap) {
        q.cr_y[base + made] = y; q.cr_xb[base + made] = xb; q.cr_xe[base + made] = x - 1; made++;
      }
      xb = x;
    }
    if (x == q.x_size - 1 && !ends && v == 1 && hh <= 1 && base + made < cap) {
      q.cr_y[base + made] = y; q.cr_xb[base + made] = xb; q.cr_xe[base + made] = q.x_size - 1; made++;
    }
    pg = g; pv = v; ps = single;
  }
}

__global__ void k_qp_mark_rows(QpArgs q) {  // markCutRowToGridKernel (PostProcessKernel.hpp:163-181)
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= q.num_rows) return;
  for (int x = q.cr_xb[r]; x <= q.cr_xe[r]; x++) atomicMax(&q.pix_row[(size_t)q.cr_y[r] * q.x_size + x], r);
}

__global__ void k_qp_cell_rows(QpArgs q) {  // markCutRowToCellKernel (PostProcessKernel.hpp:183-213)
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= q.N) return;
  int r = q.pix_row[(size_t)q.ly[c] * q.x_size + q.lx[c]];
  int rr = (q.h[c] == 1 && r >= 0) ? r : q.num_rows;
  q.cell_row[c] = rr;
  atomicAdd(&q.cr_cnt[rr], 1);
}

__global__ void k_qp_scatter(QpArgs q) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= q.N) return;
  int r = q.cell_row[c];
  int s = atomicAdd(&q.cr_fill[r], 1);
  q.cr_cells[q.cr_off[r] + s] = c;
}

// Per cut row: cells in ascending id (the stable grouping), then the
// ascending bitonic network on lx (bitonicSortAscending, PostProcessKernel.hpp:215-261)
__global__ __launch_bounds__(256) void k_qp_sort_rows(QpArgs q, int npad) {
  extern __shared__ int qlds[];
  int* key = qlds;
  int* val = qlds + npad;
  const int r = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
  const int off = q.cr_off[r], n = q.cr_off[r + 1] - off;
  int na = 1;
  while (na < n) na <<= 1;
  for (int i = t; i < na; i += nt) val[i] = i < n ? q.cr_cells[off + i] : INT_MAX;
  __syncthreads();
  for (int k = 2; k <= na; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < na; i += nt) {
        int ij = i ^ j;
        if (ij > i) {
          int a0 = val[i], a1 = val[ij];
          if (((i & k) == 0) ? (a0 > a1) : (a0 < a1)) { val[i] = a1; val[ij] = a0; }
        }
      }
      __syncthreads();
    }
  for (int i = t; i < npad; i += nt) {
    if (i < n) { key[i] = q.lx[val[i]]; }
    else { key[i] = INT_MAX; val[i] = -1; }
  }
  __syncthreads();
  for (int k = 2; k <= npad; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < npad; i += nt) {
        int ij = i ^ j;
        if (ij < i) {
          int ki = key[i], kj = key[ij];
          bool sw = ((i & k) == 0) ? (ki < kj) : (ki > kj);
          if (sw) {
            key[i] = kj; key[ij] = ki;
            int v0 = val[i]; val[i] = val[ij]; val[ij] = v0;
          }
        }
      }
      __syncthreads();
    }
  for (int i = t; i < n; i += nt) q.cr_cells[off + i] = val[i];
}

__device__ __forceinline__ unsigned long long dkey(double v) {  // order-preserving
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(unsigned long long k) {
  unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

// assignLocalIndexKernel + generateConstraintsKernel + precomputeCPrimeKernel
// (QpSolverKernels.hpp:96-199); one thread per cut row
__global__ void k_qp_init(QpArgs q) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= q.num_rows) return;
  const int off = q.cr_off[r], n = q.cr_off[r + 1] - off;
  const int rlx = q.site_width * q.cr_xb[r] + q.core_lx;
  const int rux = q.site_width * (q.cr_xe[r] + 1) + q.core_lx;
  const double rho = q.rho;
  for (int i = 0; i < n; i++) {
    int g = q.cr_cells[off + i];
    q.x[off + i] = double(q.core_lx + q.site_width * q.lx[g]);
    q.x0[off + i] = double(q.ox[g]);
    q.cp[off + i] = (i == 0) ? -rho / (1 + 2 * rho) : -rho / (1 + 2 * rho + rho * q.cp[off + i - 1]);
  }
  for (int i = 0; i < n + 1; i++) {
    int cid = off + r + i;
    int pw = (i == 0) ? 0 : q.site_width * q.w[q.cr_cells[off + i - 1]];
    q.bound[cid] = (n == 0) ? 0.0 : (i == 0) ? double(rlx) : (i == n) ? double(pw - rux) : double(pw);
    q.slack[cid] = 0.0;
    q.dualv[cid] = 0.0;
  }
}

// One ADMM iteration of CudaQpSolver::solve (CudaQpSolver.cu:242-267) for one
// cut row: updatePrimalXKernel, runThomasKernel, computeConstraintValueKernel,
// slack_update_functor, dual_update_functor, then min(Bx - b).
__global__ void k_qp_iter(QpArgs q) {
  if (q.qctl[2]) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  double mn = __longlong_as_double(0x7ff0000000000000ll);  // +inf
  if (r < q.num_rows) {
    const int off = q.cr_off[r], n = q.cr_off[r + 1] - off;
    const double rho = q.rho;
    const int cb = off + r;  // first constraint of the row
    for (int i = 0; i < n; i++) {
      int a = cb + i, b = cb + i + 1;  // pos / neg constraint of local cell off+i
      double v = q.x0[off + i];
      v -= rho * (q.slack[a] - q.bound[a] + q.dualv[a] / rho);
      v += rho * (q.slack[b] - q.bound[b] + q.dualv[b] / rho);
      q.din[off + i] = v;
    }
    const double ta = -rho, tb = 1 + 2 * rho;
    for (int i = 0; i < n; i++) {  // thomasAlgorithm (QpSolverKernels.hpp:68-94)
      int l = off + i;
      q.dp[l] = (i == 0) ? q.din[l] / tb : (q.din[l] - ta * q.dp[l - 1]) / (tb - ta * q.cp[l - 1]);
    }
    for (int i = n - 1; i >= 0; i--) {
      int l = off + i;
      q.x[l] = (i == n - 1) ? q.dp[l] : q.dp[l] - q.cp[l] * q.x[l + 1];
    }
    for (int i = 0; i < n + 1; i++) {
      int k = cb + i;
      double pp;
      if (n == 0) pp = 0.0;
      else if (i == n) pp = -q.x[off + i - 1];
      else if (i == 0) pp = q.x[off];
      else pp = q.x[off + i] - q.x[off + i - 1];
      double sl = fmin(0.0, -pp + q.bound[k] - q.dualv[k] / rho);
      q.slack[k] = sl;
      q.dualv[k] = q.dualv[k] + rho * (pp + sl - q.bound[k]);
      mn = fmin(mn, pp - q.bound[k]);
    }
  }
  unsigned long long kk = dkey(mn);
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t = __shfl_down(kk, o);
    kk = t < kk ? t : kk;
  }
  if ((threadIdx.x & 63) == 0) atomicMin(&q.qctl[0], kk);
}

// CudaQpSolver::checkTermination (CudaQpSolver.cu:118-130)
__global__ void k_qp_finish(QpArgs q) {
  if (q.qctl[2]) return;
  double mn = dkey_inv(q.qctl[0]);
  q.qctl[0] = ~0ull;
  int it = (int)q.qctl[1];
  q.qctl[3] = it + 1;
  bool conv = !(mn < -1.0 || it < q.min_iter);
  q.qctl[1] = it + 1;
  if (conv || it + 1 >= q.max_iter) q.qctl[2] = 1;
}

// siteAlignDPKernel (QpSolverKernels.hpp:358-475) / siteAlignNaiveKernel
// (QpSolverKernels.hpp:287-356), one thread per cut row
__global__ void k_qp_align(QpArgs q, int use_dp) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= q.num_rows) return;
  const int off = q.cr_off[r], n = q.cr_off[r + 1] - off;
  if (n == 0) return;
  const int sw = q.site_width, cl = q.core_lx;
  const int rlx = sw * q.cr_xb[r] + cl, rux = sw * (q.cr_xe[r] + 1) + cl;
  const double INF = __longlong_as_double(0x7ff0000000000000ll);
  if (use_dp) {
    double cL = 0, cR = 0;
    int uxl = rlx, uxr = rlx;
    for (int i = 0; i < n; i++) {
      int l = off + i, g = q.cr_cells[l];
      int wd = q.w[g] * sw;
      int orig = (int)q.x0[l];
      int clamp = max(rlx, (int)q.x[l]);
      int dist = clamp - cl;
      int nl = (dist / sw) * sw + cl, nr = (dist / sw + 1) * sw + cl;
      double dl = abs(nl - orig), dr = abs(nr - orig);
      double ml, mr;
      if (i == 0) {
        ml = dl + ((nl < rlx) ? INF : 0.0);
        mr = dr + ((nr < rlx || nr + wd > rux) ? INF : 0.0);
        q.bl[l] = -1; q.br[l] = -1;
      } else {
        double a1 = cL + dl + ((nl < uxl) ? INF : 0.0);
        double a2 = cR + dl + ((nl < uxr) ? INF : 0.0);
        if (a1 <= a2) { ml = a1; q.bl[l] = 0; } else { ml = a2; q.bl[l] = 1; }
        double bp = (nr + wd > rux) ? INF : 0.0;
        double b1 = cL + dr + (((nr < uxl) ? INF : 0.0) + bp);
        double b2 = cR + dr + (((nr < uxr) ? INF : 0.0) + bp);
        if (b1 <= b2) { mr = b1; q.br[l] = 0; } else { mr = b2; q.br[l] = 1; }
      }
      cL = ml; cR = mr;
      uxl = nl + wd; uxr = nr + wd;
    }
    int st = (cL <= cR) ? 0 : 1;
    for (int i = n - 1; i >= 0; i--) {
      int l = off + i, g = q.cr_cells[l];
      int clamp = max(rlx, (int)q.x[l]);
      int dist = clamp - cl;
      q.lx[g] = (st == 0) ? dist / sw : dist / sw + 1;
      st = (st == 0) ? q.bl[l] : q.br[l];
    }
  } else {
    int prev = rlx;
    for (int i = 0; i < n; i++) {
      int l = off + i, g = q.cr_cells[l];
      int wd = q.w[g] * sw;
      int orig = (int)q.x0[l];
      int clamp = max(rlx, (int)q.x[l]);
      int dist = clamp - cl;
      int nl = (dist / sw) * sw + cl, nr = (dist / sw + 1) * sw + cl;
      double dl = abs(nl - orig), dr = abs(nr - orig);
      int nx = nl;
      if (!(dl <= dr)) {
        if (i < n - 1) {
          int nd = (int)q.x[l + 1] - cl;
          int nll = (nd / sw) * sw + cl;
          if (nll >= nr + wd) nx = nr;
        }
        if (i == n - 1 && nr + wd <= rux) nx = nr;
      }
      if (prev > nx) nx = nr;
      prev = nx + wd;
      q.lx[g] = (nx - cl) / sw;
      q.x[l] = double(nx);
    }
  }
}

}  // namespace lgdev

int lg_qp_refine(LgHandle* h, hipStream_t s) {
  Args& a = h->a;
  const int N = a.N;
  if (N == 0) return LG_OK;
  QpArgs q{};
  q.x_size = a.x_size; q.y_size = a.y_size; q.site_width = a.site_width; q.core_lx = a.core_lx; q.N = N;
  q.pcode = a.pcode; q.h = a.h; q.w = a.w; q.ox = a.ox; q.lx = a.lx; q.ly = a.ly;
  q.rho = (double)h->p.qp_admm_rho;
  q.max_iter = h->p.qp_admm_max_iter;
  q.min_iter = h->p.qp_admm_min_iter;
  std::vector<void*> tmp;
  auto talloc = [&](auto** p, size_t n) -> int {
    int st = halloc_tmp(h, p, n);
    if (st == LG_OK) tmp.push_back((void*)*p);
    return st;
  };
  auto cleanup = [&](int code) {
    for (void* b : tmp) hfree_tmp(h, b);
    return code;
  };
  const size_t P = h->P;
  int32_t *stamp, *row_count, *row_off, *pix_row, *cell_row;
  int st;
  if ((st = talloc(&stamp, P)) || (st = talloc(&row_count, a.y_size)) || (st = talloc(&row_off, a.y_size + 1)) ||
      (st = talloc(&pix_row, P)) || (st = talloc(&cell_row, N)))
    return cleanup(st);
  q.stamp = stamp; q.row_count = row_count; q.row_off = row_off; q.pix_row = pix_row; q.cell_row = cell_row;
  if (hipMemsetAsync(stamp, 0xff, P * sizeof(int32_t), s) != hipSuccess ||
      hipMemsetAsync(pix_row, 0xff, P * sizeof(int32_t), s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp memset"));
  const int nb = (N + 255) / 256, nby = (a.y_size + 255) / 256;
  hipLaunchKernelGGL(k_qp_stamp, dim3(nb), dim3(256), 0, s, q);
  hipLaunchKernelGGL(k_qp_count_rows, dim3(nby), dim3(256), 0, s, q);
  std::vector<int32_t> cnt(a.y_size), roff(a.y_size + 1, 0);
  if (hipMemcpyAsync(cnt.data(), row_count, a.y_size * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp count rows"));
  for (int y = 0; y < a.y_size; y++) roff[y + 1] = roff[y] + cnt[y];
  const int nr = roff[a.y_size];
  q.num_rows = nr;
  if (hipMemcpyAsync(row_off, roff.data(), roff.size() * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp offsets"));
  int32_t *cy, *cxb, *cxe, *ccnt, *coff, *cfill, *ccells;
  if ((st = talloc(&cy, nr)) || (st = talloc(&cxb, nr)) || (st = talloc(&cxe, nr)) || (st = talloc(&ccnt, nr + 1)) ||
      (st = talloc(&coff, nr + 2)) || (st = talloc(&cfill, nr + 1)) || (st = talloc(&ccells, N)))
    return cleanup(st);
  q.cr_y = cy; q.cr_xb = cxb; q.cr_xe = cxe; q.cr_cnt = ccnt; q.cr_off = coff; q.cr_fill = cfill; q.cr_cells = ccells;
  // slots counted but not generated stay (y, xb, xe) = (0, 0, 0), as the
  // reference's zero-initialised vectors
  (void)hipMemsetAsync(cy, 0, std::max(nr, 1) * sizeof(int32_t), s);
  (void)hipMemsetAsync(cxb, 0, std::max(nr, 1) * sizeof(int32_t), s);
  (void)hipMemsetAsync(cxe, 0, std::max(nr, 1) * sizeof(int32_t), s);
  (void)hipMemsetAsync(ccnt, 0, (nr + 1) * sizeof(int32_t), s);
  (void)hipMemsetAsync(cfill, 0, (nr + 1) * sizeof(int32_t), s);
  hipLaunchKernelGGL(k_qp_gen_rows, dim3(nby), dim3(256), 0, s, q);
  if (nr > 0) hipLaunchKernelGGL(k_qp_mark_rows, dim3((nr + 255) / 256), dim3(256), 0, s, q);
  hipLaunchKernelGGL(k_qp_cell_rows, dim3(nb), dim3(256), 0, s, q);
  std::vector<int32_t> rc(nr + 1), ro(nr + 2, 0);
  if (hipMemcpyAsync(rc.data(), ccnt, (nr + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp cell rows"));
  int mx = 0;
  for (int r = 0; r <= nr; r++) ro[r + 1] = ro[r] + rc[r];
  for (int r = 0; r < nr; r++) mx = std::max(mx, rc[r]);
  const int nm = ro[nr];
  q.num_local = nm;
  if (hipMemcpyAsync(coff, ro.data(), ro.size() * sizeof(int32_t), hipMemcpyHostToDevice, s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp row offsets"));
  hipLaunchKernelGGL(k_qp_scatter, dim3(nb), dim3(256), 0, s, q);
  int npad = 1;
  while (npad < mx) npad <<= 1;
  if ((size_t)npad * 8 > 160 * 1024) return cleanup(fail(LG_ERR_STATE, "qp: too many cells in one cut row"));
  if (nr > 0) hipLaunchKernelGGL(k_qp_sort_rows, dim3(nr), dim3(256), (size_t)npad * 8, s, q, npad);
  double *x, *x0, *cp, *din, *dp, *bound, *slack, *dualv;
  int32_t *bl, *br;
  unsigned long long* qctl;
  const size_t nc = (size_t)nm + nr;
  if ((st = talloc(&x, nm)) || (st = talloc(&x0, nm)) || (st = talloc(&cp, nm)) || (st = talloc(&din, nm)) ||
      (st = talloc(&dp, nm)) || (st = talloc(&bound, nc)) || (st = talloc(&slack, nc)) || (st = talloc(&dualv, nc)) ||
      (st = talloc(&bl, nm)) || (st = talloc(&br, nm)) || (st = talloc(&qctl, 4)))
    return cleanup(st);
  q.x = x; q.x0 = x0; q.cp = cp; q.din = din; q.dp = dp; q.bound = bound; q.slack = slack; q.dualv = dualv;
  q.bl = bl; q.br = br; q.qctl = qctl;
  unsigned long long qinit[4] = {~0ull, 0, 0, 0};
  if (hipMemcpyAsync(qctl, qinit, sizeof(qinit), hipMemcpyHostToDevice, s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp ctl"));
  const int nbr = (nr + 63) / 64;
  if (nr > 0) {
    hipLaunchKernelGGL(k_qp_init, dim3(nbr), dim3(64), 0, s, q);
    if (q.max_iter > 0) {
      int batch = 16;
      unsigned long long hq[4];
      for (;;) {
        for (int b = 0; b < batch; b++) {
          hipLaunchKernelGGL(k_qp_iter, dim3(nbr), dim3(64), 0, s, q);
          hipLaunchKernelGGL(k_qp_finish, dim3(1), dim3(1), 0, s, q);
        }
        if (hipMemcpyAsync(hq, qctl, sizeof(hq), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
          return cleanup(fail(LG_ERR_DEVICE, "qp iterate"));
        if (hq[2]) break;
        batch = std::min(batch * 2, 256);
      }
    }
    hipLaunchKernelGGL(k_qp_align, dim3(nbr), dim3(64), 0, s, q, h->p.qp_admm_dp_site_align);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return cleanup(fail(LG_ERR_DEVICE, "qp refine kernels"));
  return cleanup(LG_OK);
}
