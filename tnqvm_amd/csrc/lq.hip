// lq.hip -- blocked Householder LQ (row-oriented) used to precondition the one-sided Jacobi SVD.
//
// For X (r x c, r <= c):  X = L Q~  (L lower triangular, Q~ orthonormal rows), then
// L^H = L_Z Q_Z and the Jacobi runs on the rows of R2 = L_Z^H.  X = Q_Z^H R2 Q~, so X and R2
// share their singular values and the right singular vectors of X are those of R2 mapped
// through Q~; the Jacobi's final rows Y_J (sigma_k v_k^H) become [Y_J | 0] Q~.  The two QR
// steps make R2 close to diagonal with a graded diagonal, and the Jacobi needs about a third
// fewer sweeps at 512 x 512 (Drmac & Veselic, SIAM J. Matrix Anal. Appl. 29 (2008) 1322,
// there with column pivoting; unpivoted here, which measured the same on MPS matrices).
//
// Reflectors act on rows from the right: y -> y - 2 <y, w> w with <y, w> = sum_l y_l conj(w_l)
// and w a unit row vector.  A panel of 16 reflectors w_k0 .. w_k0+15 composes to
// H_k0 ... H_k0+15 = I - W^H T W (W the 16 reflector rows, T upper triangular, T_jj = 2,
// T_{0:j,j} = -2 T_{0:j,0:j} G_{0:j,j}, G_ij = <w_i, w_j>); its inverse is I - W^H T^H W.
// Panels of 16 rows: 32-row panels halve the passes over the trailing matrix but put twice as
// long a serial factorization on every launch's critical path, and measured slower.
#include "kernels.h"
#include "device_util.h"

namespace hipmps {

constexpr int LQB = 16;  // rows per panel (reflectors per block reflector)
constexpr int LQR = 16;  // rows per MFMA tile in the block-reflector application

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double2 wave_sum2(double2 v) { return make_double2(wave_sum(v.x), wave_sum(v.y)); }

// Element w_k[col] of reflector k of M (support col >= k; head in W0, tail in M's row k).
__device__ __forceinline__ double2 refl(const double2* M, size_t ld, const double2* W0, int k, int col) {
  if (col < k) return make_double2(0, 0);
  if (col == k) return W0[k];
  return M[(size_t)k * ld + col];
}

// Factor rows [k0, k0 + 16) of T.M over columns [k0, cols) and form the panel's T factor.
// Whole workgroup (256 threads); rows stay in memory (L2), each wave updates every fourth row.
__device__ void lq_factor_panel(const LqTask& T, int k0, double2 (*sG)[LQB + 1], double2 (*sT)[LQB + 1]) {
  double2* M = T.M;
  const size_t ld = T.ld;
  const int C = T.cols;
  const int k1 = min(k0 + LQB, T.rows);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int k = k0; k < k1; ++k) {
    const double2* rk = M + (size_t)k * ld;
    double s = 0.0;
    for (int l = k + 1 + lane; l < C; l += 64) s += cabs2(rk[l]);
    s = wave_sum(s);
    const double2 a = rk[k];
    const double nrm = sqrt(cabs2(a) + s);
    double2 beta = make_double2(0, 0), w0 = make_double2(0, 0);
    double scale = 0.0;
    if (nrm > 0.0) {
      const double aa = sqrt(cabs2(a));
      const double2 ph = aa > 0.0 ? cscale(a, 1.0 / aa) : make_double2(1.0, 0.0);
      beta = cscale(ph, -nrm);  // -e^{i arg a} |x|: no cancellation in a - beta
      const double2 v0 = make_double2(a.x - beta.x, a.y - beta.y);
      scale = 1.0 / sqrt(cabs2(v0) + s);
      w0 = cscale(v0, scale);
    }
    // remaining rows of the panel: y -= 2 <y, w> w
    for (int i = k + 1; i < k1; ++i) {
      if (((i - k0) & 3) != wave) continue;
      double2* yi = M + (size_t)i * ld;
      double2 d = lane == 0 ? cfmac(w0, yi[k], make_double2(0, 0)) : make_double2(0, 0);
      for (int l = k + 1 + lane; l < C; l += 64) d = cfmac(cscale(rk[l], scale), yi[l], d);
      d = wave_sum2(d);  // <y, w> = sum y_l conj(w_l)  (cfmac(a, b) adds conj(a) b)
      const double2 m2d = cscale(d, -2.0);
      if (lane == 0) yi[k] = cfma(m2d, w0, yi[k]);
      for (int l = k + 1 + lane; l < C; l += 64) yi[l] = cfma(m2d, cscale(rk[l], scale), yi[l]);
    }
    __syncthreads();  // every read of row k done
    for (int l = k + 1 + tid; l < C; l += 256) M[(size_t)k * ld + l] = cscale(rk[l], scale);
    if (tid == 0) {
      M[(size_t)k * ld + k] = beta;
      T.W0[k] = w0;
    }
    __syncthreads();
  }
  // G_ij = <w_i, w_j> for i < j (common support starts at column k0 + j)
  for (int i = 0; i < LQB; ++i) {
    if ((i & 3) != wave) continue;
    for (int j = i + 1; j < LQB; ++j) {
      double2 d = make_double2(0, 0);
      if (k0 + j < k1)
        for (int l = k0 + j + lane; l < C; l += 64)
          d = cfmac(refl(M, ld, T.W0, k0 + j, l), refl(M, ld, T.W0, k0 + i, l), d);
      d = wave_sum2(d);
      if (lane == 0) sG[i][j] = d;
    }
  }
  __syncthreads();
  // T (upper triangular), column by column: T_jj = 2 (0 past the last row of the panel),
  // T_ij = -2 sum_{i <= m < j} T_im G_mj for i < j
  for (int j = 0; j < LQB; ++j) {
    if (tid < LQB) {
      double2 v = make_double2(0, 0);
      if (tid < j) {
        for (int m = tid; m < j; ++m) v = cfma(sT[tid][m], sG[m][j], v);
        v = cscale(v, -2.0);
      } else if (tid == j) {
        v = make_double2(k0 + j < k1 ? 2.0 : 0.0, 0.0);
      }
      sT[tid][j] = v;
    }
    __syncthreads();
  }
  double2* Tp = T.T + (size_t)(k0 / LQB) * LQB * LQB;
  for (int e = tid; e < LQB * LQB; e += 256) Tp[e] = sT[e / LQB][e % LQB];
  __syncthreads();
}

// Rows [r0, r0 + 16) (clipped to r1) of Y (ld ly): Y <- Y (I - W^H T' W) over columns [k0, C),
// T' = T (conjT = 0) or T^H (conjT = 1), W the LQB reflector rows of the panel at k0.
// P = Y W^H (16 x LQB) by MFMA with the column range split over the 4 waves and the next
// 4-column unit prefetched, Q = P T' with T' staged in LDS, then Y -= Q W by MFMA, each wave
// owning whole 16-column tiles.
__device__ void lq_apply_panel(const LqTask& T, int k0, double2* Y, size_t ly, int r0, int r1, int conjT,
                               double2 (*sP)[LQB + 1], double2 (*sQ)[LQB + 1], double2 (*sT)[LQB + 1]) {
  const double2* M = T.M;
  const size_t ld = T.ld;
  const int C = T.cols;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, k4 = lane >> 4;
  const int k1 = min(k0 + LQB, T.rows);
  const bool rowok = r0 + l16 < r1;
  const double2* yrow = Y + (size_t)(r0 + l16) * ly;
  constexpr int NG = LQB / LQR;  // 16-column groups of P
  auto wload = [&](int j, int col) -> double2 {
    return (j < k1 && col < C) ? refl(M, ld, T.W0, j, col) : make_double2(0, 0);
  };
  // ---- P[i][j] = sum_c Y[r0+i][c] conj(w_j[c]); B operand of group g: reflector k0 + 16g + l16 ----
  mfma_d4 pr[NG], pim[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    pr[g] = mfma_d4{0, 0, 0, 0};
    pim[g] = mfma_d4{0, 0, 0, 0};
  }
  const int nu = (C - k0 + 3) >> 2;
  int u = wave;
  int col = k0 + 4 * u + k4;
  double2 a = (u < nu && rowok && col < C) ? yrow[col] : make_double2(0, 0);
  double2 wv[NG];
#pragma unroll
  for (int g = 0; g < NG; ++g) wv[g] = u < nu ? wload(k0 + LQR * g + l16, col) : make_double2(0, 0);
  for (; u < nu; u += 4) {
    const double2 x = a;
    double2 y[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) y[g] = wv[g];
    const int ncol = col + 16;
    const bool nok = u + 4 < nu;
    a = (nok && rowok && ncol < C) ? yrow[ncol] : make_double2(0, 0);
#pragma unroll
    for (int g = 0; g < NG; ++g) wv[g] = nok ? wload(k0 + LQR * g + l16, ncol) : make_double2(0, 0);
    col = ncol;
    // x conj(y): re = x.x y.x + x.y y.y, im = x.y y.x - x.x y.y
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      pr[g] = mfma64(x.x, y[g].x, pr[g]);
      pr[g] = mfma64(x.y, y[g].y, pr[g]);
      pim[g] = mfma64(x.y, y[g].x, pim[g]);
      pim[g] = mfma64(-x.x, y[g].y, pim[g]);
    }
  }
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double2 e = make_double2(pr[g][v], pim[g][v]);
          if (w) e = cadd(e, sP[4 * v + k4][LQR * g + l16]);
          sP[4 * v + k4][LQR * g + l16] = e;
        }
    }
    __syncthreads();
  }
  // ---- Q = P T' (16 x LQB), T' staged 16 rows at a time in sT ----
  {
    const double2* Tp = T.T + (size_t)(k0 / LQB) * LQB * LQB;
    double2 q[NG];
    const int i = tid >> 4, j0 = tid & 15;
#pragma unroll
    for (int h = 0; h < NG; ++h) {  // rows 16h .. 16h + 15 of T'
      for (int e = tid; e < LQR * LQB; e += 256) {
        const int m = e / LQB, j = e % LQB, mm = LQR * h + m;
        sT[m][j] = conjT ? cconj(Tp[j * LQB + mm]) : Tp[mm * LQB + j];
      }
      __syncthreads();
#pragma unroll
      for (int x = 0; x < NG; ++x) {
        const int j = j0 + LQR * x;
        double2 acc = h ? q[x] : make_double2(0, 0);
#pragma unroll 4
        for (int m = 0; m < LQR; ++m) acc = cfma(sP[i][LQR * h + m], sT[m][j], acc);
        q[x] = acc;
      }
      __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < NG; ++x) sQ[i][LQR * x + j0] = q[x];
  }
  __syncthreads();
  // ---- Y -= Q W over 16-column tiles ----
  double qr[LQB / 4], qi[LQB / 4];
#pragma unroll
  for (int kk = 0; kk < LQB / 4; ++kk) {
    const double2 q = sQ[l16][kk * 4 + k4];  // A[i][k] = Q[i][j], j = 4kk + k
    qr[kk] = q.x;
    qi[kk] = q.y;
  }
  const int nt = (C - k0 + 15) >> 4;
  for (int t = wave; t < nt; t += 4) {
    const int c0 = k0 + 16 * t + l16;
    mfma_d4 dr = {0, 0, 0, 0}, di = {0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < LQB / 4; ++kk) {
      const double2 w = wload(k0 + kk * 4 + k4, c0);  // B[k][c] = w_j[c]
      dr = mfma64(qr[kk], w.x, dr);
      dr = mfma64(-qi[kk], w.y, dr);
      di = mfma64(qr[kk], w.y, di);
      di = mfma64(qi[kk], w.x, di);
    }
    if (c0 < C) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = r0 + 4 * v + k4;
        if (row < r1) {
          double2* y = Y + (size_t)row * ly + c0;
          *y = make_double2(y->x - dr[v], y->y - di[v]);
        }
      }
    }
  }
  __syncthreads();  // sP / sQ are reused by the caller
}

__global__ __launch_bounds__(256) void k_lq_panel0(const LqTask* tasks) {
  __shared__ double2 sG[LQB][LQB + 1];
  __shared__ double2 sT[LQB][LQB + 1];
  const LqTask& T = tasks[blockIdx.x];
  if (T.rows <= 0) return;
  lq_factor_panel(T, 0, sG, sT);
}

// Trailing update with panel p (k1 = end of panel p): workgroup (0, task) owns the rows of
// panel p + 1, [k1, k1 + LQB), updates them and then factors them; workgroup (x > 0, task)
// owns rows [k1 + LQB + 16 (x - 1), + 16).
__global__ __launch_bounds__(256) void k_lq_update(const LqTask* tasks, int p) {
  __shared__ double2 sP[LQB][LQB + 1];
  __shared__ double2 sQ[LQB][LQB + 1];
  __shared__ double2 sT[LQR][LQB + 1];
  const LqTask& T = tasks[blockIdx.y];
  const int k0 = p * LQB;
  if (k0 >= T.rows) return;
  const int k1 = min(k0 + LQB, T.rows);
  if (blockIdx.x == 0) {
    if (k1 >= T.rows) return;
    const int e1 = min(k1 + LQB, T.rows);
    for (int r0 = k1; r0 < e1; r0 += LQR) lq_apply_panel(T, k0, T.M, T.ld, r0, e1, 0, sP, sQ, sT);
    lq_factor_panel(T, k1, sP, sQ);
    return;
  }
  const int r0 = k1 + LQB + LQR * ((int)blockIdx.x - 1);
  if (r0 >= T.rows) return;
  lq_apply_panel(T, k0, T.M, T.ld, r0, min(r0 + LQR, T.rows), 0, sP, sQ, sT);
}

// Tgt <- Tgt (I - W_p^H T_p^H W_p) on rows [16x, 16x + 16) of Tgt.
__global__ __launch_bounds__(256) void k_lq_apply_q(const LqTask* tasks, int p) {
  __shared__ double2 sP[LQB][LQB + 1];
  __shared__ double2 sQ[LQB][LQB + 1];
  __shared__ double2 sT[LQR][LQB + 1];
  const LqTask& T = tasks[blockIdx.y];
  const int k0 = p * LQB;
  if (k0 >= T.rows) return;
  const int r0 = LQR * (int)blockIdx.x;
  if (r0 >= T.rows_t) return;
  lq_apply_panel(T, k0, T.Tgt, T.ld_t, r0, min(r0 + LQR, T.rows_t), 1, sP, sQ, sT);
}

// dst[i][j] = conj(src[j][i]) for j >= i and i < n, 0 otherwise (the conjugate transpose of
// src's lower triangle, an upper-triangular n x n matrix); rows n .. rows_dst - 1 are zero.
__global__ __launch_bounds__(256) void k_lower_to_upper_h(const TriTask* tasks) {
  __shared__ double2 tile[32][33];
  const TriTask& T = tasks[blockIdx.z];
  const int n = T.n;
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  if (i0 >= T.rows_dst || j0 >= n) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int y = ty; y < 32; y += 8) {  // src rows j0 + y, columns i0 + tx
    const int r = j0 + y, c = i0 + tx;
    tile[y][tx] = (r < n && c < n && r >= c) ? T.src[(size_t)r * T.ld_src + c] : make_double2(0, 0);
  }
  __syncthreads();
  for (int y = ty; y < 32; y += 8) {
    const int i = i0 + y, j = j0 + tx;
    if (i < T.rows_dst && j < n) T.dst[(size_t)i * T.ld_dst + j] = cconj(tile[tx][y]);
  }
}

__global__ __launch_bounds__(256) void k_widen(const TriTask* tasks) {
  const TriTask& T = tasks[blockIdx.y];
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t total = (size_t)T.rows_dst * T.ld_dst;
  if (e >= total) return;
  const int i = (int)(e / T.ld_dst), j = (int)(e % T.ld_dst);
  T.dst[e] = j < T.n ? T.src[(size_t)i * T.ld_src + j] : make_double2(0, 0);
}

// ---- factorisation with the panel in registers, look-ahead on two streams ----------------------
// Factor rows [k0, k1) of T.M (panel p) over columns [k0, cols) with the panel in LDS (16 rows of
// at most LQ_REG_MAX_COLS columns, 128 KiB): the same reflectors as lq_factor_panel, with every
// dependent access of the 16 serial steps in LDS instead of memory.  Eight waves; wave w updates
// the rows i = w (mod 8), its inner products reduced within the wave.
__global__ __launch_bounds__(512) void k_lq_factor_lds(const LqTask* tasks, int p) {
  constexpr int NT = 512, NW = NT / 64;
  __shared__ double2 sy[LQB][LQ_REG_MAX_COLS];
  __shared__ double2 sG[LQB][LQB + 1];
  __shared__ double2 sT[LQB][LQB + 1];
  __shared__ double2 s_w0[LQB];
  __shared__ double s_scale[LQB];
  const LqTask& T = tasks[blockIdx.x];
  const int k0 = p * LQB;
  if (k0 >= T.rows) return;
  const int k1 = min(k0 + LQB, T.rows), nr = k1 - k0;
  const int W = T.cols - k0;  // columns k0 .. cols - 1 (local index c = column - k0)
  double2* M = T.M;
  const size_t ld = T.ld;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  for (int e = tid; e < LQB * W; e += NT) {
    const int i = e / W, c = e % W;
    sy[i][c] = i < nr ? M[(size_t)(k0 + i) * ld + k0 + c] : make_double2(0, 0);
  }
  __syncthreads();
  // Reflector of row k (wave 0 only): w = (w0 at column k, scale * row k beyond), beta onto the
  // diagonal.  Row k's tail stays unscaled in LDS: every use multiplies by s_scale[k], the same
  // single rounding as storing it scaled.
  auto reflector = [&](int k) {
    double s = 0.0;
    for (int c = k + 1 + lane; c < W; c += 64) s += cabs2(sy[k][c]);
    s = wave_sum(s);
    if (lane == 0) {
      const double2 a = sy[k][k];
      const double nrm = sqrt(cabs2(a) + s);
      double2 beta = make_double2(0, 0), w0 = make_double2(0, 0);
      double scale = 0.0;
      if (nrm > 0.0) {
        const double aa = sqrt(cabs2(a));
        const double2 ph = aa > 0.0 ? cscale(a, 1.0 / aa) : make_double2(1.0, 0.0);
        beta = cscale(ph, -nrm);  // -e^{i arg a} |x|: no cancellation in a - beta
        const double2 v0 = make_double2(a.x - beta.x, a.y - beta.y);
        scale = 1.0 / sqrt(cabs2(v0) + s);
        w0 = cscale(v0, scale);
      }
      sy[k][k] = beta;
      s_w0[k] = w0;
      s_scale[k] = scale;
    }
  };
  if (wave == 0 && nr > 0) reflector(0);
  __syncthreads();
  // One barrier per reflector: wave w updates rows i = r + 1 + w (mod 8); wave 0 owns row r + 1,
  // the next pivot, and forms its reflector as soon as that row is updated.
  for (int r = 0; r < nr; ++r) {
    const double2 w0 = s_w0[r];
    const double scale = s_scale[r];
    for (int i = r + 1 + wave; i < nr; i += NW) {
      double2 d = lane == 0 ? cfmac(w0, sy[i][r], make_double2(0, 0)) : make_double2(0, 0);
      for (int c = r + 1 + lane; c < W; c += 64) d = cfmac(cscale(sy[r][c], scale), sy[i][c], d);
      d = wave_sum2(d);  // <y, w> = sum y_l conj(w_l)
      const double2 m2d = cscale(d, -2.0);
      if (lane == 0) sy[i][r] = cfma(m2d, w0, sy[i][r]);
      for (int c = r + 1 + lane; c < W; c += 64) sy[i][c] = cfma(m2d, cscale(sy[r][c], scale), sy[i][c]);
      if (i == r + 1) {
        // the row's new values were written by other lanes of this wave: order them before the
        // reflector's reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        reflector(i);
      }
    }
    __syncthreads();
  }
  // G_ij = <w_i, w_j> = sum_l conj(w_j[l]) w_i[l] for i < j, support from local column j on
  for (int pr = wave; pr < LQB * LQB; pr += NW) {
    const int i = pr / LQB, j = pr % LQB;
    if (j <= i) continue;
    double2 d = make_double2(0, 0);
    if (j < nr) {
      const double si = s_scale[i], sj = s_scale[j];
      if (lane == 0) d = cfmac(s_w0[j], cscale(sy[i][j], si), d);
      for (int c = j + 1 + lane; c < W; c += 64) d = cfmac(cscale(sy[j][c], sj), cscale(sy[i][c], si), d);
    }
    d = wave_sum2(d);
    if (lane == 0) sG[i][j] = d;
  }
  __syncthreads();
  // T column by column in wave 0 (lanes 0..15), ordered within the wave
  if (wave == 0) {
    for (int j = 0; j < LQB; ++j) {
      if (lane < LQB) {
        double2 t = make_double2(0, 0);
        if (lane < j) {
          for (int m = lane; m < j; ++m) t = cfma(sT[lane][m], sG[m][j], t);
          t = cscale(t, -2.0);
        } else if (lane == j) {
          t = make_double2(k0 + j < k1 ? 2.0 : 0.0, 0.0);
        }
        sT[lane][j] = t;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __syncthreads();
  double2* Tp = T.T + (size_t)p * LQB * LQB;
  for (int e = tid; e < LQB * LQB; e += NT) Tp[e] = sT[e / LQB][e % LQB];
  if (tid < nr) T.W0[k0 + tid] = s_w0[tid];
  for (int e = tid; e < nr * W; e += NT) {
    const int i = e / W, c = e % W;
    M[(size_t)(k0 + i) * ld + k0 + c] = c > i ? cscale(sy[i][c], s_scale[i]) : sy[i][c];
  }
}

// Rows of panel p + 1 (16 rows from k1) with the block reflector of panel p
__global__ __launch_bounds__(256) void k_lq_next(const LqTask* tasks, int p) {
  __shared__ double2 sP[LQB][LQB + 1];
  __shared__ double2 sQ[LQB][LQB + 1];
  __shared__ double2 sT[LQR][LQB + 1];
  const LqTask& T = tasks[blockIdx.x];
  const int k0 = p * LQB;
  const int k1 = min(k0 + LQB, T.rows);
  if (k1 >= T.rows) return;
  lq_apply_panel(T, k0, T.M, T.ld, k1, min(k1 + LQB, T.rows), 0, sP, sQ, sT);
}

// Rows beyond panel p + 1 with the block reflector of panel p, 16 rows per workgroup
__global__ __launch_bounds__(256) void k_lq_rest(const LqTask* tasks, int p) {
  __shared__ double2 sP[LQB][LQB + 1];
  __shared__ double2 sQ[LQB][LQB + 1];
  __shared__ double2 sT[LQR][LQB + 1];
  const LqTask& T = tasks[blockIdx.y];
  const int k0 = p * LQB;
  const int r0 = k0 + 2 * LQB + LQR * (int)blockIdx.x;
  if (r0 >= T.rows) return;
  lq_apply_panel(T, k0, T.M, T.ld, r0, min(r0 + LQR, T.rows), 0, sP, sQ, sT);
}

hipError_t launch_lq_lookahead(const LqTask* d_tasks, int ntasks, int max_rows, int max_cols, hipStream_t sa,
                               hipStream_t sb, hipEvent_t ea, hipEvent_t eb) {
  if (ntasks <= 0 || max_rows <= 0) return hipSuccess;
  if (max_cols > LQ_REG_MAX_COLS) return hipErrorNotSupported;
  auto factor = [&](int p) { hipLaunchKernelGGL(k_lq_factor_lds, dim3(ntasks), dim3(512), 0, sa, d_tasks, p); };
  hipError_t e = hipSuccess;
  // sb starts after what sa has queued so far (the matrices are ready there)
  if ((e = hipEventRecord(ea, sa)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(sb, ea, 0)) != hipSuccess) return e;
  factor(0);
  if ((e = hipEventRecord(ea, sa)) != hipSuccess) return e;
  const int panels = (max_rows + LQB - 1) / LQB;
  for (int p = 0; p + 1 < panels; ++p) {
    // sa: rows of panel p + 1 need the rest-update of panel p - 1 (recorded in eb), sb: the rest
    // update of panel p needs panel p's reflectors (recorded in ea); both waits are enqueued before
    // the events are recorded again below
    if (p > 0 && (e = hipStreamWaitEvent(sa, eb, 0)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(sb, ea, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_lq_next, dim3(ntasks), dim3(256), 0, sa, d_tasks, p);
    factor(p + 1);
    if ((e = hipEventRecord(ea, sa)) != hipSuccess) return e;
    const int trailing = max_rows - (p + 2) * LQB;
    if (trailing > 0)
      hipLaunchKernelGGL(k_lq_rest, dim3((trailing + LQR - 1) / LQR, ntasks), dim3(256), 0, sb, d_tasks, p);
    if ((e = hipEventRecord(eb, sb)) != hipSuccess) return e;
  }
  // everything after this on sa sees both streams' work
  if ((e = hipEventRecord(eb, sb)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(sa, eb, 0)) != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_widen(const TriTask* d_tasks, int ntasks, int max_cols, int max_rows_dst, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  const size_t total = (size_t)max_cols * max_rows_dst;
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((total + 255) / 256), ntasks), dim3(256), 0, s, d_tasks);
  return hipGetLastError();
}

hipError_t launch_lq(const LqTask* d_tasks, int ntasks, int max_rows, hipStream_t s) {
  if (ntasks <= 0 || max_rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lq_panel0, dim3(ntasks), dim3(256), 0, s, d_tasks);
  const int panels = (max_rows + LQB - 1) / LQB;
  for (int p = 0; p + 1 < panels; ++p) {
    const int trailing = max_rows - (p + 1) * LQB;  // > 0
    const int blocks = 1 + (trailing > LQB ? (trailing - LQB + LQR - 1) / LQR : 0);
    hipLaunchKernelGGL(k_lq_update, dim3(blocks, ntasks), dim3(256), 0, s, d_tasks, p);
  }
  return hipGetLastError();
}

hipError_t launch_lq_apply_q(const LqTask* d_tasks, int ntasks, int max_rows, int max_rows_t, hipStream_t s) {
  if (ntasks <= 0 || max_rows <= 0) return hipSuccess;
  const int panels = (max_rows + LQB - 1) / LQB;
  for (int p = panels - 1; p >= 0; --p)
    hipLaunchKernelGGL(k_lq_apply_q, dim3((max_rows_t + LQR - 1) / LQR, ntasks), dim3(256), 0, s, d_tasks, p);
  return hipGetLastError();
}

hipError_t launch_lower_to_upper_h(const TriTask* d_tasks, int ntasks, int max_n, int max_rows_dst, hipStream_t s) {
  if (ntasks <= 0 || max_n <= 0) return hipSuccess;
  dim3 grid((max_n + 31) / 32, (max_rows_dst + 31) / 32, ntasks);
  hipLaunchKernelGGL(k_lower_to_upper_h, grid, dim3(256), 0, s, d_tasks);
  return hipGetLastError();
}

// Rank cut of the preconditioned factor J (n x n, ld_dst; TriTask.dst, .n): rows whose squared
// norm is below 1e-28 ||J||_F^2 are set to zero before the sweeps.  Zeroing them perturbs J by
// at most 1e-14 sqrt(n) ||J||_F, which bounds the change of every singular value (Weyl); the cut
// is made only where that bound is below 1e-3 of sigma_cut[t], the singular value under which
// the truncation drops an index, so neither a kept value nor the bond dimension moves.  Rows of
// rounding noise (a rank-deficient theta', as on purified sites) then no longer keep the
// relative convergence test rotating; a full-rank J is only read.
__global__ __launch_bounds__(256) void k_rank_cut(const TriTask* tasks, const double* sigma_cut) {
  __shared__ double rn[4096];
  __shared__ double red[4];
  __shared__ double s_thr;
  const TriTask& T = tasks[blockIdx.x];
  const double cut = sigma_cut[blockIdx.x];
  if (!(cut > 0.0)) return;
  const int n = T.n;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  double part = 0.0;
  for (int i = wave; i < n; i += 4) {
    const double2* row = T.dst + (size_t)i * T.ld_dst;
    double a = 0.0;
    for (int j = lane; j < n; j += 64) a += cabs2(row[j]);
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) rn[i] = a;
    part += a;
  }
  if (lane == 0) red[wave] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double fro2 = red[0] + red[1] + red[2] + red[3];
    s_thr = (1e-14 * sqrt((double)n * fro2) <= 1e-3 * cut) ? 1e-28 * fro2 : -1.0;
  }
  __syncthreads();
  const double thr = s_thr;
  if (thr < 0.0) return;
  for (int i = wave; i < n; i += 4) {
    if (rn[i] >= thr) continue;
    double2* row = T.dst + (size_t)i * T.ld_dst;
    for (int j = lane; j < n; j += 64) row[j] = make_double2(0.0, 0.0);
  }
}

hipError_t launch_rank_cut(const TriTask* d_tasks, const double* d_sigma_cut, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rank_cut, dim3(ntasks), dim3(256), 0, s, d_tasks, d_sigma_cut);
  return hipGetLastError();
}

}  // namespace hipmps
