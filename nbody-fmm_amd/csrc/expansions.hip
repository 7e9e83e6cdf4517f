// expansions.hip -- solid-harmonic expansions of the gfx950 FMM.
//
// Storage.  A cell's multipole M and local L are kept in the real layout of
// the reference's PowerSet<3> (power_set3D.cpp:49-63): K = (p+1)^2 doubles,
// index n^2 -> Re X_n^0, n^2+2m-1 -> Re X_n^m, n^2+2m -> Im X_n^m (m >= 1),
// padded to KP (a multiple of 16).  Coefficients are Schmidt-normalised and
// scaled by the cell scale rho_l (the largest cell edge at level l):
//   M_n^m = sum_i q_i conj(R_n^m((x_i - c)/rho))
//   phi(x) = (1/rho) sum_{n,m} L_n^m R_n^m((x - c)/rho)
// where R_n^m are the regular harmonics of PowerSet<3> (power_set3D.cpp:10-47).
// With this scaling the M2L operator of an offset class d (integer cell
// offset in [-3,3]^3) is the same matrix at every level, so M2L is one dense
// KP x KP contraction per class, executed with v_mfma_f64_16x16x4f64.
//
// Passes.  The upward pass forms M directly from the particles of every cell
// that is an M2L source (P2M at every level, the particles of a level-l cell
// are contiguous in Morton order); the downward pass evaluates the local of
// every M2L target directly at its particles (L2P at every level).  This is
// algebraically the M2M/L2L chain of the reference (node.cpp:14-43): M2M of
// exact multipoles is exact and L2L of a degree-p local is exact, so the two
// differ only by rounding, while the cost follows the (few) cells that take
// part in M2L instead of every cell of the tree.
#include "nbfmm_internal.h"

#include <cmath>
#include <cstring>

namespace nbfmm {
namespace {

constexpr int kT = 256;
constexpr int P2M_CHUNK = 1024;

// Coefficient tables (host-computed once per order, uploaded to c->consts).
struct Consts {
  double mmf[kMaxOrder + 2];                  // -sqrt((2k-1)/(2k))
  double s2m1[kMaxOrder + 2];                 // sqrt(2m+1)
  double ra[kMaxOrder + 2][kMaxOrder + 2];    // (2n-1)/sqrt(n^2-m^2)
  double rb[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n-1)^2-m^2)/sqrt(n^2-m^2)
  double cz[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n+m)(n-m))
  double cu[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n-m+1)(n-m))
  double cd[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n+m+1)(n+m))
};

void fill_consts(Consts &C) {
  std::memset(&C, 0, sizeof(C));
  for (int k = 1; k <= kMaxOrder + 1; ++k) C.mmf[k] = -std::sqrt((2.0 * k - 1.0) / (2.0 * k));
  for (int m = 0; m <= kMaxOrder + 1; ++m) C.s2m1[m] = std::sqrt(2.0 * m + 1.0);
  for (int n = 0; n <= kMaxOrder + 1; ++n)
    for (int m = 0; m <= n && m <= kMaxOrder + 1; ++m) {
      if (n >= 2 && m <= n - 2) {
        double den = std::sqrt((double)(n * n - m * m));
        C.ra[n][m] = (2.0 * n - 1.0) / den;
        C.rb[n][m] = std::sqrt((double)((n - 1) * (n - 1) - m * m)) / den;
      }
      C.cz[n][m] = std::sqrt((double)((n + m) * (n - m)));
      C.cu[n][m] = std::sqrt((double)((n - m + 1) * (n - m)));
      C.cd[n][m] = std::sqrt((double)((n + m + 1) * (n + m)));
    }
}


__device__ __forceinline__ void cell_centre(uint64_t key, int lev, const LevelGeo &g, double c[3]) {
  uint64_t k = key;
  uint32_t q[3] = {0, 0, 0};
  for (int b = 0; b < lev; ++b) {
    q[0] |= (uint32_t)((k >> (3 * b)) & 1u) << b;
    q[1] |= (uint32_t)((k >> (3 * b + 1)) & 1u) << b;
    q[2] |= (uint32_t)((k >> (3 * b + 2)) & 1u) << b;
  }
  for (int d = 0; d < 3; ++d) c[d] = g.lo[d] + ((double)q[d] + 0.5) * g.h[lev][d];
}

__device__ __forceinline__ int level_of(int32_t cell, const LevelGeo &g) {
  int l = 0;
  while (l < g.depth && cell >= g.lvl_start[l + 1]) ++l;
  return l;
}

// ---- M2L operators -----------------------------------------------------------------
// Yoshida singular harmonics S_N^M (N <= 2p, 0 <= M <= N) of the scaled class
// offset, then the real KP x KP matrix in MFMA A-fragment order:
//   frag[(rb*KS + ks)*64 + lane] = T[rb*16 + lane%16][ks*4 + lane/16].
// L_j^k = (-1)^{j+k}/N_jk sum_{n,m} (M_n^m/N_nm) S_{n+j}^{m-k}(D),  N_nm = sqrt((n+m)!(n-m)!)
__global__ void __launch_bounds__(kT) k_build_ops(int p, int KP, double rx, double ry, double rz,
                                                  double *__restrict__ frag) {
  const int cls = blockIdx.x;
  const int dx = cls % 7 - 3, dy = (cls / 7) % 7 - 3, dz = cls / 49 - 3;
  const int P2 = 2 * p;
  __shared__ double Sre[(2 * kMaxOrder + 1) * (2 * kMaxOrder + 2) / 2 + 8];
  __shared__ double Sim[(2 * kMaxOrder + 1) * (2 * kMaxOrder + 2) / 2 + 8];
  __shared__ double lf[2 * kMaxOrder + 2]; // log-free factorials as doubles
  const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
  if (threadIdx.x <= (unsigned)(2 * kMaxOrder + 1)) {
    double f = 1;
    for (int k = 2; k <= (int)threadIdx.x; ++k) f *= k;
    lf[threadIdx.x] = f;
  }
  // the class offset d is source - target; M2L needs c_L - c_M = -d
  const double X = -dx * rx, Y = -dy * ry, Z = -dz * rz;
  const double r2 = X * X + Y * Y + Z * Z;
  if (!near && (int)threadIdx.x <= P2) {
    // column M = threadIdx.x of S (Yoshida normalisation)
    const int M = threadIdx.x;
    const double ir2 = 1.0 / r2;
    double sr = 1.0 / sqrt(r2), si = 0;
    for (int k = 1; k <= M; ++k) { // S_k^k = -(2k-1)(x+iy)/r^2 S_{k-1}^{k-1}
      double f = -(2.0 * k - 1.0) * ir2;
      double nr = f * (X * sr - Y * si), ni = f * (X * si + Y * sr);
      sr = nr;
      si = ni;
    }
    int idx = M * (M + 1) / 2 + M;
    Sre[idx] = sr;
    Sim[idx] = si;
    double pr = 0, pi = 0, cr = sr, ci = si;
    for (int N = M + 1; N <= P2; ++N) {
      double a = (2.0 * N - 1.0) * Z * ir2;
      double b = (double)((N - 1) * (N - 1) - M * M) * ir2;
      double nr = a * cr - b * pr, ni = a * ci - b * pi;
      pr = cr;
      pi = ci;
      cr = nr;
      ci = ni;
      int id = N * (N + 1) / 2 + M;
      Sre[id] = cr;
      Sim[id] = ci;
    }
  }
  __syncthreads();
  const int K = (p + 1) * (p + 1);
  const int KS = KP / 4;
  double *out = frag + (size_t)cls * KP * KP;
  for (int e = threadIdx.x; e < KP * KP; e += kT) {
    const int lane = e & 63, blk = e >> 6;
    const int rb = blk / KS, ks = blk % KS;
    const int row = rb * 16 + (lane & 15), col = ks * 4 + (lane >> 4);
    double v = 0;
    if (!near && row < K && col < K) {
      int j = (int)sqrt((double)row);
      while (j * j > row) --j;
      while ((j + 1) * (j + 1) <= row) ++j;
      int rr = row - j * j; // 0 -> (k=0, re); 2k-1 -> (k, re); 2k -> (k, im)
      int k = (rr + 1) / 2, oim = (rr > 0 && (rr & 1) == 0);
      int n = (int)sqrt((double)col);
      while (n * n > col) --n;
      while ((n + 1) * (n + 1) <= col) ++n;
      int cc = col - n * n;
      int m = (cc + 1) / 2, iim = (cc > 0 && (cc & 1) == 0);
      // S_{n+j}^{M'} with sign/conjugation for negative M'
      auto S = [&](int NN, int MM, double &re, double &im) {
        if (MM >= 0) {
          re = Sre[NN * (NN + 1) / 2 + MM];
          im = Sim[NN * (NN + 1) / 2 + MM];
        } else {
          re = Sre[NN * (NN + 1) / 2 - MM];
          im = -Sim[NN * (NN + 1) / 2 - MM];
          if ((-MM) & 1) {
            re = -re;
            im = -im;
          }
        }
      };
      const double Njk = sqrt(lf[j + k] * lf[j - k]);
      const double Nnm = sqrt(lf[n + m] * lf[n - m]);
      const double cjk = (((j + k) & 1) ? -1.0 : 1.0) / Njk / Nnm;
      double Are, Aim;
      S(n + j, m - k, Are, Aim);
      if (m == 0) {
        v = cjk * (oim ? Aim : Are);
      } else {
        double Bre, Bim;
        S(n + j, m + k, Bre, Bim);
        Bim = -Bim; // conj
        if (k & 1) {
          Bre = -Bre;
          Bim = -Bim;
        }
        // term = a (A+B) + i b (A-B), M = a + i b
        if (!oim)
          v = iim ? -cjk * (Aim - Bim) : cjk * (Are + Bre);
        else
          v = iim ? cjk * (Are - Bre) : cjk * (Aim + Bim);
      }
    }
    out[e] = v;
  }
}

// ---- P2M ---------------------------------------------------------------------------
// Item = (source slot, particle range).  Lane = particle; the orders m are
// dealt round-robin to the 4 waves of the block, so every wave runs one
// column m at a time with no divergence: q R_m^m by the sectoral recurrence,
// then R_n^m, n = m+1..p, by the three-term recurrence in n, accumulated per
// lane; one xor-shuffle reduction per coefficient at the end of the column.
// T is the coefficient type (double, or float in mixed precision); the
// offset from the cell centre is formed in FP64 and scaled before rounding.
template <class T>
__global__ void __launch_bounds__(kT) k_p2m(int p, int KP, const int4 *__restrict__ items, const double4 *__restrict__ pos,
                                            const int32_t *__restrict__ src_cells, const uint64_t *__restrict__ ckey,
                                            LevelGeo g, const Consts *__restrict__ C, T *__restrict__ Mexp) {
  const int4 it = items[blockIdx.x];
  const int slot = it.x, beg = it.y, end = it.z;
  const int32_t cell = src_cells[slot];
  const int lev = level_of(cell, g);
  double ctr[3];
  cell_centre(ckey[cell], lev, g, ctr);
  const double ir = g.inv_rho[lev];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T *M = Mexp + (size_t)slot * KP;
  for (int m = wave; m <= p; m += kT / 64) {
    T ar[kMaxOrder + 1], ai[kMaxOrder + 1];
#pragma unroll
    for (int j = 0; j <= kMaxOrder; ++j) {
      ar[j] = 0;
      ai[j] = 0;
    }
    const int nj = p - m; // j = n - m in [0, nj]
    for (int i = beg + lane; i < end; i += 64) {
      const double4 q = pos[i];
      const T x = (T)((q.x - ctr[0]) * ir), y = (T)((q.y - ctr[1]) * ir), z = (T)((q.z - ctr[2]) * ir);
      const T r2 = x * x + y * y + z * z;
      T vr = (T)q.w, vi = 0; // q * R_m^m
      for (int k = 1; k <= m; ++k) {
        const T f = (T)C->mmf[k];
        const T nr = f * (x * vr - y * vi), ni = f * (x * vi + y * vr);
        vr = nr;
        vi = ni;
      }
      ar[0] += vr;
      ai[0] -= vi;
      T pr = vr, pi = vi;
      T cr = 0, ci = 0;
      if (nj >= 1) {
        const T f = z * (T)C->s2m1[m];
        cr = f * pr;
        ci = f * pi;
        ar[1] += cr;
        ai[1] -= ci;
      }
#pragma unroll
      for (int j = 2; j <= kMaxOrder; ++j) {
        if (j <= nj) {
          const int n = m + j;
          const T a = (T)C->ra[n][m] * z, b = (T)C->rb[n][m] * r2;
          const T nr = a * cr - b * pr, ni = a * ci - b * pi;
          pr = cr;
          pi = ci;
          cr = nr;
          ci = ni;
          ar[j] += cr;
          ai[j] -= ci;
        }
      }
    }
#pragma unroll
    for (int j = 0; j <= kMaxOrder; ++j) {
      if (j <= nj) {
        T vr = ar[j], vi = ai[j];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
          vr += __shfl_xor(vr, o);
          vi += __shfl_xor(vi, o);
        }
        if (lane == 0) {
          const int n = m + j;
          if (m == 0) {
            atomicAdd(&M[n * n], vr);
          } else {
            atomicAdd(&M[n * n + 2 * m - 1], vr);
            atomicAdd(&M[n * n + 2 * m], vi);
          }
        }
      }
    }
  }
}

// ---- L2P --------------------------------------------------------------------------------
// grad phi at every particle from the locals of its ancestors that are M2L
// targets:  grad = (1/rho^2) sum_{n'<p, m'>=0} w(m') Re[R_n'^m'(t) G(n'+1, m')]
// with G_z = cz L_n^m', G_x = (cu L_n^{m'-1} - cd L_n^{m'+1})/2,
// G_y = -(i/2)(cu L_n^{m'-1} + cd L_n^{m'+1}), L_n^{-1} = -conj(L_n^1).
template <class T>
struct cpx {
  T x, y;
};
template <class T>
__device__ __forceinline__ cpx<T> lget(const T *L, int n, int m, int p) {
  if (m > n || n > p) return {0, 0};
  if (m == 0) return {L[n * n], 0};
  if (m > 0) return {L[n * n + 2 * m - 1], L[n * n + 2 * m]};
  // m == -1
  return {-L[n * n + 1], L[n * n + 2]};
}

// T as in k_p2m; the gradient is accumulated over levels in FP64.
template <class T>
__global__ void __launch_bounds__(kT) k_l2p(int p, int KP, int64_t n, const double4 *__restrict__ pos,
                                            const int32_t *__restrict__ pcell, const int32_t *__restrict__ cparent,
                                            const uint64_t *__restrict__ ckey, const int32_t *__restrict__ tgt_slot,
                                            LevelGeo g, const Consts *__restrict__ C, const T *__restrict__ Lexp,
                                            double *__restrict__ grad) {
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double4 q = pos[i];
  double gx = 0, gy = 0, gz = 0;
  int32_t cell = pcell[i];
  for (int lev = g.depth; lev >= 2; --lev) {
    const int32_t ts = tgt_slot[cell];
    if (ts >= 0) {
      double ctr[3];
      cell_centre(ckey[cell], lev, g, ctr);
      const double ir = g.inv_rho[lev];
      const T x = (T)((q.x - ctr[0]) * ir), y = (T)((q.y - ctr[1]) * ir), z = (T)((q.z - ctr[2]) * ir);
      const T r2 = x * x + y * y + z * z;
      const T *L = Lexp + (size_t)ts * KP;
      T hx = 0, hy = 0, hz = 0;
      T vr = 1, vi = 0; // R_m^m
      for (int m = 0; m < p; ++m) {
        if (m > 0) {
          const T f = (T)C->mmf[m];
          const T nr = f * (x * vr - y * vi), ni = f * (x * vi + y * vr);
          vr = nr;
          vi = ni;
        }
        const T w = m == 0 ? (T)1 : (T)2;
        T pr = 0, pi = 0, cr = vr, ci = vi;
        for (int np = m; np < p; ++np) {
          if (np == m + 1) {
            pr = cr;
            pi = ci;
            const T f = z * (T)C->s2m1[m];
            cr = f * pr;
            ci = f * pi;
          } else if (np > m + 1) {
            const T a = (T)C->ra[np][m] * z, b = (T)C->rb[np][m] * r2;
            const T nr = a * cr - b * pr, ni = a * ci - b * pi;
            pr = cr;
            pi = ci;
            cr = nr;
            ci = ni;
          }
          const int nn = np + 1;
          const cpx<T> Lm = lget(L, nn, m - 1, p), L0 = lget(L, nn, m, p), Lp = lget(L, nn, m + 1, p);
          const T cz = (T)C->cz[nn][m], cu = (T)C->cu[nn][m], cdd = (T)C->cd[nn][m];
          // Re(R * G)
          hz += w * cz * (cr * L0.x - ci * L0.y);
          const T ux = cu * Lm.x - cdd * Lp.x, uy = cu * Lm.y - cdd * Lp.y;
          hx += (T)0.5 * w * (cr * ux - ci * uy);
          const T vx = cu * Lm.x + cdd * Lp.x, vy = cu * Lm.y + cdd * Lp.y;
          hy += (T)0.5 * w * (cr * vy + ci * vx);
        }
      }
      const double s = ir * ir;
      gx += s * (double)hx;
      gy += s * (double)hy;
      gz += s * (double)hz;
    }
    cell = cparent[cell];
  }
  grad[3 * i] = gx;
  grad[3 * i + 1] = gy;
  grad[3 * i + 2] = gz;
}

__global__ void k_p2m_count(const int32_t *src_cells, int64_t nsrc, const int32_t *cbeg, const int32_t *cend,
                            int32_t *cnt) {
  int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= nsrc) return;
  int32_t c = src_cells[s];
  cnt[s] = (cend[c] - cbeg[c] + P2M_CHUNK - 1) / P2M_CHUNK;
}
__global__ void k_p2m_items(const int32_t *src_cells, int64_t nsrc, const int32_t *cbeg, const int32_t *cend,
                            const int32_t *off, int4 *items) {
  int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= nsrc) return;
  int32_t c = src_cells[s];
  int32_t w = off[s];
  for (int32_t b = cbeg[c]; b < cend[c]; b += P2M_CHUNK)
    items[w++] = make_int4((int)s, b, min(b + P2M_CHUNK, cend[c]), 0);
}
__global__ void k_terms(const int32_t *cells, int64_t ncell, const int32_t *cbeg, const int32_t *cend,
                        unsigned long long *out) {
  __shared__ unsigned long long s[kT];
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  s[threadIdx.x] = i < ncell ? (unsigned long long)(cend[cells[i]] - cbeg[cells[i]]) : 0ull;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT > 0 ? (n + kT - 1) / kT : 1); }

} // namespace

int padded_k(int p) { return ((p + 1) * (p + 1) + 15) / 16 * 16; }

void run_expansions(nbfmm_ctx *c) {
  Plan &P = c->plan;
  hipStream_t s = c->stream;
  const int p = (int)c->prm.order;
  const int KP = padded_k(p);
  P.kp = KP;
  const int D = P.geo.depth;
  const bool f32 = c->prm.precision == NBFMM_MIXED;
  LevelGeo g;
  std::memset(&g, 0, sizeof(g));
  for (int d = 0; d < 3; ++d) g.lo[d] = P.geo.lo[d];
  for (int l = 0; l <= D; ++l) {
    for (int d = 0; d < 3; ++d) g.h[l][d] = P.geo.h(l, d);
    g.inv_rho[l] = 1.0 / P.geo.rho(l);
  }
  for (int l = 0; l <= D + 1; ++l) g.lvl_start[l] = P.lvl_start[l];
  g.depth = D;
  // constants
  Consts *dC = (Consts *)dbuf_ensure(c->consts, sizeof(Consts));
  {
    static Consts hC;
    static bool init = false;
    if (!init) {
      fill_consts(hC);
      init = true;
    }
    NB_HIP(hipMemcpyAsync(dC, &hC, sizeof(Consts), hipMemcpyHostToDevice, s));
  }
  double *grad = dget<double>(c->grad, (size_t)P.n * 3);
  if (P.n_tgt == 0) {
    NB_HIP(hipMemsetAsync(grad, 0, sizeof(double) * 3 * (size_t)P.n, s));
    P.n_p2m_terms = P.n_l2p_terms = 0;
    mark(c, 3);
    mark(c, 4);
    mark(c, 5);
    return;
  }
  const int32_t *cbeg = (const int32_t *)c->cbeg.p, *cend = (const int32_t *)c->cend.p;
  const uint64_t *ckey = (const uint64_t *)c->ckey.p;
  // ---- P2M of every M2L source cell ----
  int32_t *pcnt = dget<int32_t>(c->p2m_off, (size_t)P.n_src * 2 + 2 + 8);
  int32_t *poff = pcnt + P.n_src + 1;
  hipLaunchKernelGGL(k_p2m_count, dim3(nblk(P.n_src)), dim3(kT), 0, s, (const int32_t *)c->src_cells.p, P.n_src,
                     cbeg, cend, pcnt);
  NB_LAUNCH();
  NB_HIP(hipMemsetAsync(pcnt + P.n_src, 0, sizeof(int32_t), s));
  scan_excl_i32(pcnt, poff, P.n_src + 1, c->scan_tmp, s);
  int32_t nitems;
  NB_HIP(hipMemcpyAsync(&nitems, poff + P.n_src, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  unsigned long long *terms = (unsigned long long *)((double *)c->red.p + 6 * 1024 + 48);
  NB_HIP(hipMemsetAsync(terms, 0, 2 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_terms, dim3(nblk(P.n_src)), dim3(kT), 0, s, (const int32_t *)c->src_cells.p, P.n_src, cbeg,
                     cend, terms);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_terms, dim3(nblk(P.n_tgt)), dim3(kT), 0, s, (const int32_t *)c->tgt_cells.p, P.n_tgt, cbeg,
                     cend, terms + 1);
  NB_LAUNCH();
  unsigned long long th[2];
  NB_HIP(hipMemcpyAsync(th, terms, sizeof(th), hipMemcpyDeviceToHost, s));
  NB_HIP(hipStreamSynchronize(s));
  P.n_p2m_items = nitems;
  P.n_p2m_terms = (int64_t)th[0];
  P.n_l2p_terms = (int64_t)th[1];
  int4 *items = dget<int4>(c->p2m_items, (size_t)nitems + 1);
  hipLaunchKernelGGL(k_p2m_items, dim3(nblk(P.n_src)), dim3(kT), 0, s, (const int32_t *)c->src_cells.p, P.n_src, cbeg,
                     cend, poff, items);
  NB_LAUNCH();
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  void *Mexp = dbuf_ensure(c->mexp, (size_t)P.n_src * KP * esz + 16);
  NB_HIP(hipMemsetAsync(Mexp, 0, esz * (size_t)P.n_src * KP, s));
  if (f32)
    hipLaunchKernelGGL(k_p2m<float>, dim3(nitems), dim3(kT), 0, s, p, KP, items, (const double4 *)c->pos.p,
                       (const int32_t *)c->src_cells.p, ckey, g, dC, (float *)Mexp);
  else
    hipLaunchKernelGGL(k_p2m<double>, dim3(nitems), dim3(kT), 0, s, p, KP, items, (const double4 *)c->pos.p,
                       (const int32_t *)c->src_cells.p, ckey, g, dC, (double *)Mexp);
  NB_LAUNCH();
  mark(c, 3);
  // ---- M2L ----
  // offset classes scaled by (hx, hy, hz) / rho (level independent)
  const double rho = P.geo.rho(D);
  double *frag = dget<double>(c->tmat, (size_t)kNumClasses * KP * KP);
  hipLaunchKernelGGL(k_build_ops, dim3(kNumClasses), dim3(kT), 0, s, p, KP, P.geo.h(D, 0) / rho,
                     P.geo.h(D, 1) / rho, P.geo.h(D, 2) / rho, frag);
  NB_LAUNCH();
  run_m2l(c, KP, f32);
  mark(c, 4);
  // ---- L2P ----
  if (f32)
    hipLaunchKernelGGL(k_l2p<float>, dim3(nblk(P.n)), dim3(kT), 0, s, p, KP, P.n, (const double4 *)c->pos.p,
                       (const int32_t *)c->pcell.p, (const int32_t *)c->cparent.p, ckey, (const int32_t *)c->tgt_slot.p,
                       g, dC, (const float *)c->lexp.p, grad);
  else
    hipLaunchKernelGGL(k_l2p<double>, dim3(nblk(P.n)), dim3(kT), 0, s, p, KP, P.n, (const double4 *)c->pos.p,
                       (const int32_t *)c->pcell.p, (const int32_t *)c->cparent.p, ckey, (const int32_t *)c->tgt_slot.p,
                       g, dC, (const double *)c->lexp.p, grad);
  NB_LAUNCH();
  mark(c, 5);
}

} // namespace nbfmm
