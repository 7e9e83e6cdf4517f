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
// Passes.  Upward: every cell that needs a multipole (an M2L source, or a
// child of a cell that forms its own by M2M) forms it either by P2M from its
// particles (at most kTransCut particles, or a leaf; the particles of a cell
// are contiguous in Morton order) or by M2M from its children (larger cells),
// the reference's upward pass (node.cpp:14-24) restricted to the cells that
// need it.  Downward: every cell that holds a local (an M2L target, or a
// child of a cell that pushes) either pushes it to its children by L2L
// (larger cells, node.cpp:33-43) or has it evaluated at its particles (L2P).
// M2M and L2L are MFMA contractions with one operator per child octant
// (m2l.hip); P2M and L2P run on the vector ALUs.
#include "nbfmm_internal.h"

#include <cstddef>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nbfmm {
namespace {

constexpr int kT = 256;

// Coefficient tables (host-computed once per order, uploaded to c->consts).
struct Consts {
  double mmf[kMaxOrder + 2];                  // -sqrt((2k-1)/(2k))
  double s2m1[kMaxOrder + 2];                 // sqrt(2m+1)
  double ra[kMaxOrder + 2][kMaxOrder + 2];    // (2n-1)/sqrt(n^2-m^2)
  double rb[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n-1)^2-m^2)/sqrt(n^2-m^2)
  double cz[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n+m)(n-m))
  double cu[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n-m+1)(n-m))
  double cd[kMaxOrder + 2][kMaxOrder + 2];    // sqrt((n+m+1)(n+m))
  double smm[kMaxOrder + 2];                  // mmf[1] ... mmf[m] (1 for m = 0)
};

void fill_consts(Consts &C) {
  std::memset(&C, 0, sizeof(C));
  for (int k = 1; k <= kMaxOrder + 1; ++k) C.mmf[k] = -std::sqrt((2.0 * k - 1.0) / (2.0 * k));
  for (int m = 0; m <= kMaxOrder + 1; ++m) C.s2m1[m] = std::sqrt(2.0 * m + 1.0);
  C.smm[0] = 1.0;
  for (int m = 1; m <= kMaxOrder + 1; ++m) C.smm[m] = C.smm[m - 1] * C.mmf[m];
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


// The operators (M2L classes, M2M/L2L octants, their FP32 and bf16 forms)
// depend only on the leaf geometry dp->hs, the order and the precision: they
// are rebuilt only when one of these changed since the last build (or the
// host moved a buffer: force).  Sets dp->ops_valid, which the builders read.
__global__ void k_ops_check(DevPlan *__restrict__ dp, OpsKey *__restrict__ key, int p, int KP, int f32, int force) {
  if (threadIdx.x != 0 || plan_dead(dp)) return;
  const bool same = !force && key->hs[0] == dp->hs[0] && key->hs[1] == dp->hs[1] && key->hs[2] == dp->hs[2] &&
                    key->p == p && key->kp == KP && key->f32 == f32;
  dp->ops_valid = same ? 1 : 0;
  if (!same) {
    for (int i = 0; i < 3; ++i) key->hs[i] = dp->hs[i];
    key->p = p;
    key->kp = KP;
    key->f32 = f32;
  }
}

// ---- M2L operators -----------------------------------------------------------------
// Yoshida singular harmonics S_N^M (N <= 2p, 0 <= M <= N) of the scaled class
// offset, then the real KP x KP matrix in MFMA A-fragment order:
//   frag[(rb*KS + ks)*64 + lane] = T[rb*16 + lane%16][ks*4 + lane/16].
// L_j^k = (-1)^{j+k}/N_jk sum_{n,m} (M_n^m/N_nm) S_{n+j}^{m-k}(D),  N_nm = sqrt((n+m)!(n-m)!)
// (rx, ry, rz): the leaf cell's edges over rho, from the plan
__global__ void __launch_bounds__(kT) k_build_ops(int p, int KP, const DevPlan *__restrict__ dp,
                                                  double *__restrict__ frag) {
  if (dp->ops_valid) return;
  const double rx = dp->hs[0], ry = dp->hs[1], rz = dp->hs[2];
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

// ---- M2M and L2L operators ---------------------------------------------------------------
// Child octant o (bit d set: upper half along axis d) sits at d = child centre
// - parent centre = (+-h_x, +-h_y, +-h_z) / 4 of the parent cell.  In the
// rho-scaled Schmidt storage (rho_child = rho_parent / 2) the translations are
// the same matrices at every level; with Yoshida harmonics Rt of
// dh = d / rho_parent and N_nm = sqrt((n+m)!(n-m)!):
//   M2M  M_n^m(parent) = sum_{n'<=n,k} N_nm/N_n'k 2^-n' conj(Rt_{n-n'}^{m-k}(dh)) M_n'^k(child)
//   L2L  L_j^k(child)  = sum_{n>=j,m}  N_nm/N_jk 2^-(j+1) Rt_{n-j}^{m-k}(dh) L_n^m(parent)
// (the Yoshida translations M'_n^m = sum conj(R_j^k(c - c')) M_{n-j}^{m-k} and
// L'_j^k = sum L_n^m R_{n-j}^{m-k}(c' - c) rewritten for the scaled storage).
// The real KP x KP matrix of each octant is written in MFMA A-fragment order
// like the M2L operators; negative orders follow X_n^-m = (-1)^m conj(X_n^m)
// for input and output.
__device__ __forceinline__ void yget(const double *Rre, const double *Rim, int n, int m, double &re, double &im) {
  const int a = m < 0 ? -m : m;
  re = Rre[n * (n + 1) / 2 + a];
  im = Rim[n * (n + 1) / 2 + a];
  if (m < 0) {
    im = -im;
    if (a & 1) {
      re = -re;
      im = -im;
    }
  }
}

__global__ void __launch_bounds__(kT) k_build_trans(int p, int KP, const DevPlan *__restrict__ dp,
                                                    double *__restrict__ m2m, double *__restrict__ l2l) {
  if (dp->ops_valid) return;
  const double hx = dp->hs[0], hy = dp->hs[1], hz = dp->hs[2];
  const int o = blockIdx.x, op = blockIdx.y;
  __shared__ double Rre[(kMaxOrder + 1) * (kMaxOrder + 2) / 2];
  __shared__ double Rim[(kMaxOrder + 1) * (kMaxOrder + 2) / 2];
  __shared__ double lf[2 * kMaxOrder + 2];
  if (threadIdx.x <= (unsigned)(2 * kMaxOrder + 1)) {
    double f = 1;
    for (int k = 2; k <= (int)threadIdx.x; ++k) f *= k;
    lf[threadIdx.x] = f;
  }
  if (threadIdx.x == 0) {
    const double x = (o & 1 ? 0.25 : -0.25) * hx, y = (o & 2 ? 0.25 : -0.25) * hy, z = (o & 4 ? 0.25 : -0.25) * hz;
    const double r2 = x * x + y * y + z * z;
    double sr = 1, si = 0; // Rt_m^m
    for (int m = 0; m <= p; ++m) {
      if (m > 0) {
        const double f = -0.5 / m; // Rt_m^m = -(x + iy) / (2m) Rt_{m-1}^{m-1}
        const double nr = f * (x * sr - y * si), ni = f * (x * si + y * sr);
        sr = nr;
        si = ni;
      }
      Rre[m * (m + 1) / 2 + m] = sr;
      Rim[m * (m + 1) / 2 + m] = si;
      double pr = 0, pi = 0, cr = sr, ci = si;
      for (int n = m + 1; n <= p; ++n) {
        // (n+m)(n-m) Rt_n^m = (2n-1) z Rt_{n-1}^m - r^2 Rt_{n-2}^m
        const double inv = 1.0 / ((double)(n + m) * (double)(n - m));
        const double nr = ((2.0 * n - 1.0) * z * cr - r2 * pr) * inv, ni = ((2.0 * n - 1.0) * z * ci - r2 * pi) * inv;
        pr = cr;
        pi = ci;
        cr = nr;
        ci = ni;
        Rre[n * (n + 1) / 2 + m] = cr;
        Rim[n * (n + 1) / 2 + m] = ci;
      }
    }
  }
  __syncthreads();
  const int K = (p + 1) * (p + 1), KS = KP / 4;
  auto Nf = [&](int n, int m) { return sqrt(lf[n + m] * lf[n - m]); };
  // complex coefficient of input (b, kb) in output (a, ka), any signs of ka, kb
  auto coef = [&](int a, int ka, int b, int kb, double &re, double &im) {
    re = 0;
    im = 0;
    if (op == 0) { // M2M: output (n=a, m=ka), input (n'=b, k=kb)
      if (b > a || abs(ka - kb) > a - b) return;
      yget(Rre, Rim, a - b, ka - kb, re, im);
      im = -im;
      const double f = Nf(a, ka) / Nf(b, kb) * ldexp(1.0, -b);
      re *= f;
      im *= f;
    } else { // L2L: output (j=a, k=ka), input (n=b, m=kb)
      if (b < a || abs(kb - ka) > b - a) return;
      yget(Rre, Rim, b - a, kb - ka, re, im);
      const double f = Nf(b, kb) / Nf(a, ka) * ldexp(1.0, -(a + 1));
      re *= f;
      im *= f;
    }
  };
  double *out = (op == 0 ? m2m : l2l) + (size_t)o * KP * KP;
  for (int e = threadIdx.x; e < KP * KP; e += kT) {
    const int lane = e & 63, blk = e >> 6;
    const int rb = blk / KS, ks = blk % KS;
    const int row = rb * 16 + (lane & 15), col = ks * 4 + (lane >> 4);
    double v = 0;
    if (row < K && col < K) {
      int a = (int)sqrt((double)row);
      while (a * a > row) --a;
      while ((a + 1) * (a + 1) <= row) ++a;
      const int rr = row - a * a, ka = (rr + 1) / 2, oim = rr > 0 && (rr & 1) == 0;
      int b = (int)sqrt((double)col);
      while (b * b > col) --b;
      while ((b + 1) * (b + 1) <= col) ++b;
      const int cc = col - b * b, kb = (cc + 1) / 2, iim = cc > 0 && (cc & 1) == 0;
      double pr, pi;
      coef(a, ka, b, kb, pr, pi);
      if (kb == 0) {
        v = oim ? pi : pr; // real input
      } else {
        double mr, mi;
        coef(a, ka, b, -kb, mr, mi);
        const double s = (kb & 1) ? -1.0 : 1.0;
        // input a + ib:  out = (K+ + s K-) a + i (K+ - s K-) b
        if (!iim)
          v = oim ? pi + s * mi : pr + s * mr;
        else
          v = oim ? pr - s * mr : -(pi - s * mi);
      }
    }
    out[e] = v;
  }
}

// ---- P2M ---------------------------------------------------------------------------
// One wavefront per P2M cell, four cells per workgroup.  Lanes are LP column
// lanes times 64 / LP particle slots.  Column lane c holds two columns, the
// orders m1 = c and m2 = p - c (one when they meet), so every lane has p + 2
// terms: q R_m^m = q S_m w^m (w = x + i y, S_m the product of the sectoral
// recurrence factors, w^m by binary powers), then R_n^m, n = m+1..p, by the
// three-term recurrence in n.  A slot runs every (64/LP)-th particle of the
// cell, accumulating per lane; the slots meet in xor-shuffles and the column
// lanes store the coefficients.  (The previous layout, one column per lane,
// ran p sectoral and p recurrence steps for each particle of 4 slots.)
// PM bounds the order (register arrays); P2M cells hold at most kTransCut
// particles unless they are leaves (more only where particles coincide at
// the deepest level) and mostly tens.  T is the coefficient type (double, or
// float in mixed precision); the offset from the cell centre is formed in
// FP64 and scaled before rounding.
template <class T, int LP, int PM>
__global__ void __launch_bounds__(kT) k_p2m(int p, int KP, const DevPlan *__restrict__ dp,
                                            const int32_t *__restrict__ cells,
                                            const double4 *__restrict__ pos, const int32_t *__restrict__ cbeg,
                                            const int32_t *__restrict__ cend, const int32_t *__restrict__ mslot,
                                            const uint64_t *__restrict__ ckey,
                                            const Consts *__restrict__ C, T *__restrict__ Mexp) {
  if (plan_dead(dp)) return;
  const LevelGeo &g = dp->g;
  const int ncell = dp->reads[1]; // cells forming their multipole by P2M
  static_assert(LP >= PM / 2 + 1 && 64 % LP == 0, "a column lane per pair of orders");
  constexpr int NS = 64 / LP;      // particle slots
  constexpr int N1 = PM + 1;       // terms of a lane's first column, at most
  constexpr int N2 = (PM + 1) / 2; // of its second
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k = (int)blockIdx.x * (kT / 64) + wave;
  if (k >= ncell) return; // a whole wavefront; nothing below synchronises the workgroup
  const int32_t cell = cells[k];
  const int beg = cbeg[cell], end = cend[cell];
  const int slot = mslot[cell];
  const int lev = level_of(cell, g);
  double ctr[3];
  cell_centre(ckey[cell], lev, g, ctr);
  const double ir = g.inv_rho[lev];
  const int cl = lane & (LP - 1), ps = lane / LP;
  const int m1 = cl, m2 = p - cl;
  const int n1 = 2 * cl <= p ? p - m1 + 1 : 0; // terms j = 0..p-m1 of column m1
  const int n2 = 2 * cl < p ? p - m2 + 1 : 0;  // terms of column m2 (none when m2 <= m1)
  // term j >= 1 of a column m: c_j = (a_j z) c_{j-1} - (b_j r^2) c_{j-2}
  T a1[N1], b1[N1], a2[N2], b2[N2];
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    a1[j] = j == 1 && j < n1 ? (T)C->s2m1[m1] : j >= 2 && j < n1 ? (T)C->ra[m1 + j][m1] : (T)0;
    b1[j] = j >= 2 && j < n1 ? (T)C->rb[m1 + j][m1] : (T)0;
  }
#pragma unroll
  for (int j = 0; j < N2; ++j) {
    a2[j] = j == 1 && j < n2 ? (T)C->s2m1[m2] : j >= 2 && j < n2 ? (T)C->ra[m2 + j][m2] : (T)0;
    b2[j] = j >= 2 && j < n2 ? (T)C->rb[m2 + j][m2] : (T)0;
  }
  const T s1 = n1 > 0 ? (T)C->smm[m1] : (T)0, s2 = n2 > 0 ? (T)C->smm[m2] : (T)0;
  T r1[N1], i1[N1], r2[N2], i2[N2];
#pragma unroll
  for (int j = 0; j < N1; ++j) r1[j] = i1[j] = 0;
#pragma unroll
  for (int j = 0; j < N2; ++j) r2[j] = i2[j] = 0;
  for (int i = beg + ps; i < end; i += NS) {
    const double4 q = pos[i];
    const T x = (T)((q.x - ctr[0]) * ir), y = (T)((q.y - ctr[1]) * ir), z = (T)((q.z - ctr[2]) * ir);
    const T rr = x * x + y * y + z * z;
    // w^m1 and w^m2 from w, w^2, w^4, ...
    T wr = x, wi = y, e1r = 1, e1i = 0, e2r = 1, e2i = 0;
#pragma unroll
    for (int bit = 0; (1 << bit) <= PM; ++bit) {
      if ((1 << bit) > p) break; // uniform
      if ((m1 >> bit) & 1) {
        const T tr = e1r * wr - e1i * wi;
        e1i = e1r * wi + e1i * wr;
        e1r = tr;
      }
      if ((m2 >> bit) & 1) {
        const T tr = e2r * wr - e2i * wi;
        e2i = e2r * wi + e2i * wr;
        e2r = tr;
      }
      const T sr = wr * wr - wi * wi;
      wi = 2 * wr * wi;
      wr = sr;
    }
    const T qm = (T)q.w;
    {
      T cr = qm * s1 * e1r, ci = qm * s1 * e1i, pr = 0, pi = 0; // q R_m1^m1
      r1[0] += cr;
      i1[0] -= ci;
#pragma unroll
      for (int j = 1; j < N1; ++j) {
        if (j <= p) { // uniform
          const T a = a1[j] * z, b = b1[j] * rr;
          const T nr = a * cr - b * pr, ni = a * ci - b * pi;
          pr = cr;
          pi = ci;
          cr = nr;
          ci = ni;
          r1[j] += cr;
          i1[j] -= ci;
        }
      }
    }
    {
      T cr = qm * s2 * e2r, ci = qm * s2 * e2i, pr = 0, pi = 0; // q R_m2^m2
      r2[0] += cr;
      i2[0] -= ci;
#pragma unroll
      for (int j = 1; j < N2; ++j) {
        if (2 * j <= p) { // uniform: n2 <= (p + 1) / 2
          const T a = a2[j] * z, b = b2[j] * rr;
          const T nr = a * cr - b * pr, ni = a * ci - b * pi;
          pr = cr;
          pi = ci;
          cr = nr;
          ci = ni;
          r2[j] += cr;
          i2[j] -= ci;
        }
      }
    }
  }
  // the slots of a column lane: lanes cl, cl + LP, cl + 2 LP, ...
  T *M = Mexp + (size_t)slot * KP;
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    if (j <= p) { // uniform: every lane shuffles
      T vr = r1[j], vi = i1[j];
#pragma unroll
      for (int o = LP; o < 64; o <<= 1) {
        vr += __shfl_xor(vr, o);
        vi += __shfl_xor(vi, o);
      }
      if (lane < LP && j < n1) {
        const int n = m1 + j;
        const int er = m1 == 0 ? n * n : n * n + 2 * m1 - 1;
        M[er] = vr;
        if (m1 > 0) M[er + 1] = vi;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < N2; ++j) {
    if (2 * j <= p) { // uniform
      T vr = r2[j], vi = i2[j];
#pragma unroll
      for (int o = LP; o < 64; o <<= 1) {
        vr += __shfl_xor(vr, o);
        vi += __shfl_xor(vi, o);
      }
      if (lane < LP && j < n2) {
        const int n = m2 + j;
        const int er = n * n + 2 * m2 - 1; // m2 > m1 >= 0
        M[er] = vr;
        M[er + 1] = vi;
      }
    }
  }
}

// ---- L2P --------------------------------------------------------------------------------
// grad phi at every particle from the locals of the cells on its path to the
// root that evaluate their local (eslot >= 0; the others pushed theirs down
// by L2L):  grad = (1/rho^2) sum_{n'<p, m'>=0} w(m') Re[R_n'^m'(t) G(n'+1, m')]
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

// One evaluated local at a particle (offset x, y, z, r2 = |.|^2 in cell
// units): the gradient sums hx, hy, hz and, with POT, the potential sum hv.
template <class T, bool POT>
__device__ __forceinline__ void l2p_local(const T *L, int pc, T x, T y, T z, T r2, const Consts *__restrict__ C,
                                          T &hx, T &hy, T &hz, T &hv) {
  T vr = 1, vi = 0; // R_m^m
  for (int m = 0; m < (POT ? pc + 1 : pc); ++m) {
    if (m > 0) {
      const T f = (T)C->mmf[m];
      const T nr = f * (x * vr - y * vi), ni = f * (x * vi + y * vr);
      vr = nr;
      vi = ni;
    }
    const T w = m == 0 ? (T)1 : (T)2;
    T pr = 0, pi = 0, cr = vr, ci = vi;
    for (int np = m; np < (POT ? pc + 1 : pc); ++np) {
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
      if (POT) {
        const cpx<T> Lv = lget(L, np, m, pc);
        hv += w * (cr * Lv.x - ci * Lv.y);
        if (np == pc) continue;
      }
      const int nn = np + 1;
      const cpx<T> Lm = lget(L, nn, m - 1, pc), L0 = lget(L, nn, m, pc), Lp = lget(L, nn, m + 1, pc);
      const T cz = (T)C->cz[nn][m], cu = (T)C->cu[nn][m], cdd = (T)C->cd[nn][m];
      // Re(R * G)
      hz += w * cz * (cr * L0.x - ci * L0.y);
      const T ux = cu * Lm.x - cdd * Lp.x, uy = cu * Lm.y - cdd * Lp.y;
      hx += (T)0.5 * w * (cr * ux - ci * uy);
      const T vx = cu * Lm.x + cdd * Lp.x, vy = cu * Lm.y + cdd * Lp.y;
      hy += (T)0.5 * w * (cr * vy + ci * vx);
    }
  }
}

// T as in k_p2m; the gradient is accumulated over levels in FP64.  POT also
// evaluates the potential sum (1/rho) sum_{n<=p,m>=0} w(m) Re[L_n^m R_n^m] into pot.
// When every active lane of a wavefront evaluates the same cell (consecutive
// particles in Morton order mostly share their evaluated cells), the local is
// addressed through a wave-uniform pointer, so its coefficients are read by
// scalar loads instead of one vector load per lane.
template <class T, bool POT>
__global__ void __launch_bounds__(kT) k_l2p(int p, int KP, int64_t n, const double4 *__restrict__ pos,
                                            const int32_t *__restrict__ pcell, const int32_t *__restrict__ cparent,
                                            const uint64_t *__restrict__ ckey, const int32_t *__restrict__ eslot,
                                            const DevPlan *__restrict__ dp, const Consts *__restrict__ C,
                                            const T *__restrict__ Lexp, double *__restrict__ grad,
                                            double *__restrict__ pot, const int8_t *__restrict__ ord) {
  if (plan_dead(dp)) return;
  const LevelGeo &g = dp->g;
  int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const double4 q = pos[i];
  double gx = 0, gy = 0, gz = 0, gp = 0;
  int32_t cell = pcell[i];
  for (int lev = g.depth; lev >= 2; --lev) {
    const int32_t ts = eslot[cell];
    if (ts >= 0) {
      const int pc = ord ? (int)ord[cell] : p; // order of this cell's local
      double ctr[3];
      cell_centre(ckey[cell], lev, g, ctr);
      const double ir = g.inv_rho[lev];
      const T x = (T)((q.x - ctr[0]) * ir), y = (T)((q.y - ctr[1]) * ir), z = (T)((q.z - ctr[2]) * ir);
      const T r2 = x * x + y * y + z * z;
      T hx = 0, hy = 0, hz = 0, hv = 0;
      const int32_t ts0 = __builtin_amdgcn_readfirstlane(ts);
      if (__all(ts == ts0)) {
        const int pc0 = __builtin_amdgcn_readfirstlane(pc);
        l2p_local<T, POT>(Lexp + (size_t)ts0 * KP, pc0, x, y, z, r2, C, hx, hy, hz, hv);
      } else {
        l2p_local<T, POT>(Lexp + (size_t)ts * KP, pc, x, y, z, r2, C, hx, hy, hz, hv);
      }
      const double s = ir * ir;
      gx += s * (double)hx;
      gy += s * (double)hy;
      gz += s * (double)hz;
      if (POT) gp += ir * (double)hv;
    }
    cell = cparent[cell];
  }
  grad[3 * i] = gx;
  grad[3 * i + 1] = gy;
  grad[3 * i + 2] = gz;
  if (POT) pot[i] = gp;
}

// ---- upward / downward plan ------------------------------------------------------------
// A cell needs a multipole when it is an M2L source or its parent forms its
// own by M2M; it forms it by M2M from its children when it holds more than
// kTransCut particles (and is not a leaf), else by P2M from its particles.
// A cell holds a local when it is an M2L target or its parent pushes its
// local down; it pushes (L2L) when it holds more than kTransCut particles
// (and is not a leaf), else its local is evaluated at its particles (L2P).
// Modes: 0 none, 1 P2M / L2P, 2 M2M / L2L.
constexpr int32_t kTransCut = 256;

// Modes of every cell in one pass.  A cell needs its multipole when it is an
// M2L source or its parent forms its multipole by M2M (mode 2), and mode 2
// needs a cell that needs its multipole and is big.  Parents are never leaves
// and hold their children's particles, so the ancestors of a big parent are
// all big: a cell needs its multipole iff it is a source or, with a big
// parent, some ancestor is.  Locals alike, with "has an M2L list".  Also the
// five flag arrays of stride s >= n + 1 (flags past the cells 0): M slot, P2M, M2M, L slot, L2L,
// and the particles of the P2M cells (terms[0]) and of the cells that
// evaluate their local (terms[1]); terms are zero on entry.
__global__ void __launch_bounds__(kT) k_updown_modes(const DevPlan *__restrict__ dp, int D, int64_t s,
                                                     const int32_t *__restrict__ cparent,
                                                     const int32_t *__restrict__ cbeg, const int32_t *__restrict__ cend,
                                                     const int32_t *__restrict__ is_src,
                                                     const int32_t *__restrict__ cnt_m2l, int32_t *mode_m,
                                                     int32_t *mode_l, int32_t *flags, unsigned long long *terms) {
  if (plan_dead(dp)) return;
  __shared__ unsigned long long red[2][kT / 64];
  const int64_t n = dp->nc; // cells; the flags of [n, s) are zero
  const int32_t leaf_begin = dp->lvl_start[D];
  unsigned long long tm = 0, tl = 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < s; i += (int64_t)gridDim.x * kT) {
    int mm = 0, ml = 0;
    if (i < n) {
      const bool big = i < leaf_begin && cend[i] - cbeg[i] > kTransCut;
      bool m = is_src[i] != 0, l = cnt_m2l[i] > 0;
      int32_t a = cparent[i];
      if (a >= 0 && cend[a] - cbeg[a] > kTransCut) {
        for (; a >= 0 && !(m && l); a = cparent[a]) {
          m = m || is_src[a] != 0;
          l = l || cnt_m2l[a] > 0;
        }
      }
      mm = m ? (big ? 2 : 1) : 0;
      ml = l ? (big ? 2 : 1) : 0;
      mode_m[i] = mm;
      mode_l[i] = ml;
      const unsigned long long np = (unsigned long long)(cend[i] - cbeg[i]);
      if (mm == 1) tm += np;
      if (ml == 1) tl += np;
    }
    flags[i] = mm > 0;
    flags[s + i] = mm == 1;
    flags[2 * s + i] = mm == 2;
    flags[3 * s + i] = ml > 0;
    flags[4 * s + i] = ml == 2;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    tm += __shfl_xor(tm, o);
    tl += __shfl_xor(tl, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = tm;
    red[1][threadIdx.x >> 6] = tl;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    unsigned long long t = 0;
    for (int w = 0; w < kT / 64; ++w) t += red[threadIdx.x][w];
    if (t) atomicAdd(&terms[threadIdx.x], t);
  }
}

// slots (M, L, evaluated L) and the P2M, M2M and L2L cell lists; `scan` is one
// exclusive scan over the five flag arrays, so segment k starts at scan[k s]
__global__ void k_updown_lists(const DevPlan *__restrict__ dp, int64_t s, const int32_t *__restrict__ mode_l,
                               const int32_t *__restrict__ flags, const int32_t *__restrict__ scan, int32_t *mslot,
                               int32_t *lslot, int32_t *eslot, int32_t *p2m_cells, int32_t *m2m_cells,
                               int32_t *l2l_cells) {
  if (plan_dead(dp)) return;
  const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (i >= dp->nc) return;
  mslot[i] = flags[i] ? scan[i] : -1;
  const int32_t ls = flags[3 * s + i] ? scan[3 * s + i] - scan[3 * s] : -1;
  lslot[i] = ls;
  eslot[i] = mode_l[i] == 1 ? ls : -1;
  if (flags[s + i]) p2m_cells[scan[s + i] - scan[s]] = (int32_t)i;
  if (flags[2 * s + i]) m2m_cells[scan[2 * s + i] - scan[2 * s]] = (int32_t)i;
  if (flags[4 * s + i]) l2l_cells[scan[4 * s + i] - scan[4 * s]] = (int32_t)i;
}

// The planning's sizes into the plan: the five segment totals (M cells, P2M
// cells, M2M cells, L cells, L2L cells) and the M2M / L2L cells before each
// level (the cell lists are level ordered).  With check, the capacities.
__global__ void k_up_reads(DevPlan *dp, const int32_t *__restrict__ scan, int64_t s, int D, int check, int64_t cap_m,
                           int64_t cap_p2m, int64_t cap_l, CapLevels cap_m2m, CapLevels cap_l2l) {
  const int k = threadIdx.x;
  const int32_t nc = dp->nc;
  if (k < 5) dp->reads[k] = scan[k * s + nc] - scan[k * s];
  if (k <= D + 1) {
    const int32_t ls = dp->lvl_start[k];
    dp->reads[8 + k] = scan[2 * s + ls] - scan[2 * s];
    dp->reads[32 + k] = scan[4 * s + ls] - scan[4 * s];
  }
  __syncthreads();
  if (!check || k != 0) return;
  int ovf = 0;
  if (dp->reads[0] > cap_m || dp->reads[1] > cap_p2m || dp->reads[3] > cap_l) ovf = 2;
  for (int d = 0; d < D; ++d)
    if (dp->reads[9 + d] - dp->reads[8 + d] > cap_m2m.v[d] || dp->reads[33 + d] - dp->reads[32 + d] > cap_l2l.v[d])
      ovf = 2;
  if (ovf) dp->ovf |= ovf;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT > 0 ? (n + kT - 1) / kT : 1); }

} // namespace

int padded_k(int p) { return ((p + 1) * (p + 1) + 15) / 16 * 16; }

LevelGeo level_geo(const Plan &P) {
  const int D = P.geo.depth;
  LevelGeo g;
  std::memset(&g, 0, sizeof(g));
  for (int d = 0; d < 3; ++d) g.lo[d] = P.geo.lo[d];
  for (int l = 0; l <= D; ++l) {
    for (int d = 0; d < 3; ++d) g.h[l][d] = P.geo.h(l, d);
    g.inv_rho[l] = 1.0 / P.geo.rho(l);
  }
  for (int l = 0; l <= D + 1; ++l) g.lvl_start[l] = P.lvl_start[l];
  g.depth = D;
  return g;
}

void upload_consts(nbfmm_ctx *c) {
  static Consts hC;
  static bool init = false;
  if (!init) {
    fill_consts(hC);
    init = true;
  }
  Consts *dC = (Consts *)dbuf_ensure(c->consts, sizeof(Consts));
  NB_HIP(hipMemcpy(dC, &hC, sizeof(Consts), hipMemcpyHostToDevice));
}

void run_expansions(nbfmm_ctx *c) {
  Plan &P = c->plan;
  hipStream_t s = c->stream;
  const Caps &K = c->cap;
  DevPlan *dp = dplan(c);
  int p = (int)c->prm.order; // with learned orders: the highest order in use, set below
  int KP = padded_k(p);
  P.kp = KP;
  const int D = P.geo.depth;
  const bool f32 = c->prm.precision == NBFMM_MIXED;
  const bool learned = c->prm.order_select == NBFMM_ORDER_LEARNED;
  if (K.on && learned) throw Error(NBFMM_ERR_UNSUPPORTED, "planned evaluations take a fixed order");
  P.flop_m2l = 0;
  P.m2l_stats_dev = false;
  P.mean_order = p;
  const Consts *dC = (const Consts *)c->consts.p; // uploaded at nbfmm_create
  // far-field gradient (3 per particle) and, with potentials, far-field potential sums
  const bool pot = c->phi_out != nullptr;
  double *grad = dget<double>(c->grad, (size_t)P.n * (pot ? 4 : 3));
  double *pfar = pot ? grad + 3 * (size_t)P.n : nullptr;
  if (!K.on && P.n_tgt == 0) {
    c->n_local = 0;
    P.n_p2m_cells = P.n_m2m_cells = P.n_l2l_cells = P.n_l2p_cells = 0;
    fill_zero(s, grad, sizeof(double) * (pot ? 4 : 3) * (size_t)P.n);
    P.n_p2m_terms = P.n_l2p_terms = 0;
    fork_p2p(c);
    launch_p2p(c, 0);
    launch_p2p(c, 1);
    mark(c, 14);
    mark(c, 3);
    mark(c, 13);
    mark(c, 4);
    mark(c, 5);
    return;
  }
  // the first share of P2P runs from here on (its tasks are complete), beside
  // the planning, P2M, M2M and their host reads
  fork_p2p(c);
  launch_p2p(c, 0);
  const int32_t *cbeg = (const int32_t *)c->cbeg.p, *cend = (const int32_t *)c->cend.p;
  const uint64_t *ckey = (const uint64_t *)c->ckey.p;
  const int64_t nc = K.on ? K.nc : P.n_cells; // cells, or their capacity
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  // ---- plan: modes per level (top-down), slots, lists ----
  int32_t *mode_m = dget<int32_t>(c->mode_m, (size_t)nc + 1);
  int32_t *mode_l = dget<int32_t>(c->mode_l, (size_t)nc + 1);
  const int64_t s5 = nc + 1;
  int32_t *flags = dget<int32_t>(c->flags, (size_t)5 * s5);
  int32_t *fscan = dget<int32_t>(c->fscan, (size_t)5 * s5);
  unsigned long long *terms = dp->terms;
  fill_zero(s, terms, 2 * sizeof(unsigned long long));
  // few workgroups: this runs beside the first share of P2P
  hipLaunchKernelGGL(k_updown_modes, dim3(std::min<unsigned>(512, nblk(s5))), dim3(kT), 0, s, (const DevPlan *)dp, D,
                     s5, (const int32_t *)c->cparent.p, cbeg, cend, (const int32_t *)c->is_src.p,
                     (const int32_t *)c->cnt_m2l.p, mode_m, mode_l, flags, terms);
  NB_LAUNCH();
  scan_excl_i32(flags, fscan, 5 * s5, c->scan_tmp, s);
  int32_t *mslot = dget<int32_t>(c->mslot, (size_t)nc + 1);
  int32_t *lslot = dget<int32_t>(c->lslot, (size_t)nc + 1);
  int32_t *eslot = dget<int32_t>(c->eslot, (size_t)nc + 1);
  int32_t *lists = dget<int32_t>(c->up_cells, (size_t)3 * s5);
  int32_t *p2m_cells = lists, *m2m_cells = lists + s5, *l2l_cells = lists + 2 * s5;
  hipLaunchKernelGGL(k_updown_lists, dim3(nblk(nc)), dim3(kT), 0, s, (const DevPlan *)dp, s5, (const int32_t *)mode_l,
                     (const int32_t *)flags, (const int32_t *)fscan, mslot, lslot, eslot, p2m_cells, m2m_cells,
                     l2l_cells);
  NB_LAUNCH();
  // totals and the per-level starts of the M2M and L2L lists (cells are level
  // ordered) into the plan; a direct evaluation reads them with the term counts
  {
    CapLevels cm{}, cl{};
    for (int d = 0; d <= kMaxDepth; ++d) {
      cm.v[d] = K.m2m_lev[d];
      cl.v[d] = K.l2l_lev[d];
    }
    hipLaunchKernelGGL(k_up_reads, dim3(1), dim3(64), 0, s, dp, (const int32_t *)fscan, s5, D, K.on ? 1 : 0, K.nM,
                       K.nP2M, K.nL, cm, cl);
    NB_LAUNCH();
  }
  const int32_t *h = c->hplan->reads;
  if (!K.on) {
    read_plan(c, &c->hplan->reads[0], offsetof(DevPlan, terms) + 2 * sizeof(unsigned long long) -
                                          offsetof(DevPlan, reads), s);
    c->n_local = h[3];
    P.n_p2m_cells = h[1];
    P.n_m2m_cells = h[2];        // cells forming their multipole by M2M
    P.n_l2l_cells = h[4];        // cells pushing their local by L2L
    P.n_l2p_cells = h[3] - h[4]; // cells evaluating their local at their particles
    P.n_p2m_terms = (int64_t)c->hplan->terms[0];
    P.n_l2p_terms = (int64_t)c->hplan->terms[1];
  }
  const int64_t nM = K.on ? K.nM : h[0], nP2M = K.on ? K.nP2M : h[1], nL = K.on ? K.nL : h[3];
  // M2M / L2L parents of each level: exact or the capacity
  int64_t m2m_lev[kMaxDepth + 1], l2l_lev[kMaxDepth + 1];
  for (int d = 0; d < D; ++d) {
    m2m_lev[d] = K.on ? K.m2m_lev[d] : h[9 + d] - h[8 + d];
    l2l_lev[d] = K.on ? K.l2l_lev[d] : h[33 + d] - h[32 + d];
  }
  if (!K.on) {
    c->last.nM = nM;
    c->last.nP2M = nP2M;
    c->last.nL = nL;
    for (int d = 0; d < D; ++d) {
      c->last.m2m_lev[d] = (int32_t)m2m_lev[d];
      c->last.l2l_lev[d] = (int32_t)l2l_lev[d];
    }
  }
  // orders per cell (learned selection): the expansions are formed to the
  // highest order in use
  const int8_t *ord = nullptr;
  if (learned) {
    ord = order_select(c, level_geo(P));
    p = c->ord_max;
    KP = padded_k(p);
    P.kp = KP;
    P.mean_order = p;
  }
  // ---- operators: M2L classes and the M2M / L2L octants (level independent) ----
  double *frag = dget<double>(c->tmat, (size_t)kNumClasses * KP * KP);
  double *trans = dget<double>(c->trans, (size_t)16 * KP * KP);
  {
    const void *bufs[4] = {frag, trans, nullptr, nullptr};
    if (f32) { // the FP32 and bf16 forms' buffers now, so that none moves after the check
      bufs[2] = ops_f32_buffers(c, KP, &bufs[3]);
    }
    const bool moved = std::memcmp(bufs, c->ops_bufs, sizeof(bufs)) != 0;
    OpsKey *key = dget<OpsKey>(c->ops_key, 1);
    if (moved) std::memcpy(c->ops_bufs, bufs, sizeof(bufs));
    hipLaunchKernelGGL(k_ops_check, dim3(1), dim3(64), 0, s, dp, key, p, KP, f32 ? 1 : 0, moved ? 1 : 0);
    NB_LAUNCH();
  }
  hipLaunchKernelGGL(k_build_ops, dim3(kNumClasses), dim3(kT), 0, s, p, KP, (const DevPlan *)dp, frag);
  NB_LAUNCH();
  hipLaunchKernelGGL(k_build_trans, dim3(8, 2), dim3(kT), 0, s, p, KP, (const DevPlan *)dp, trans,
                     trans + (size_t)8 * KP * KP);
  NB_LAUNCH();
  if (f32) trans_to_f32(c, KP);

  // ---- P2M (cells that form their multipole from particles) ----
  P.n_p2m_items = nP2M;
  void *Mexp = dbuf_ensure(c->mexp, (size_t)nM * KP * esz + 16);
  fill_zero(s, Mexp, esz * (size_t)nM * KP);
  if (nP2M > 0) {
    const unsigned nb = (unsigned)((nP2M + kT / 64 - 1) / (kT / 64));
#define NB_P2M(T, LP, PM)                                                                                       \
  hipLaunchKernelGGL((k_p2m<T, LP, PM>), dim3(nb), dim3(kT), 0, s, p, KP, (const DevPlan *)dp,                   \
                     (const int32_t *)p2m_cells, (const double4 *)c->pos.p, cbeg, cend, (const int32_t *)mslot,   \
                     ckey, dC, (T *)Mexp)
    if (f32) {
      if (p <= 8)
        NB_P2M(float, 8, 8);
      else if (p <= 12)
        NB_P2M(float, 8, 12);
      else if (p <= 14)
        NB_P2M(float, 8, 14);
      else
        NB_P2M(float, 16, 16);
    } else {
      if (p <= 8)
        NB_P2M(double, 8, 8);
      else if (p <= 12)
        NB_P2M(double, 8, 12);
      else if (p <= 14)
        NB_P2M(double, 8, 14);
      else
        NB_P2M(double, 16, 16);
    }
#undef NB_P2M
    NB_LAUNCH();
  }
  // ---- M2M, finest level first ----
  // in mixed precision the parents of levels < lc translate in FP64 (on
  // FP64 copies of the multipoles, rounded back to FP32 for M2L)
  const int lc = f32 ? std::min(D, c->prm.mixed_fp64_levels) : 0;
  mark(c, 14);
  for (int d = D - 1; d >= lc; --d) run_trans(c, false, KP, f32, m2m_cells, &dp->reads[8 + d], m2m_lev[d]);
  if (lc > 0) {
    double *M64 = dget<double>(c->mexp64, (size_t)nM * KP);
    widen(c, (const float *)Mexp, (int64_t)nM * KP, M64);
    for (int d = lc - 1; d >= 0; --d) run_trans64(c, false, KP, m2m_cells, &dp->reads[8 + d], m2m_lev[d], M64);
    narrow(c, M64, (int64_t)nM * KP, (float *)Mexp);
  }
  mark(c, 3);
  // ---- M2L (writes the locals of its targets) ----
  // The rest of P2P starts after the first pass's M2L contraction
  // (run_m2l), so that the contraction, which holds a CU's whole LDS per
  // workgroup, is not starved by it, and the memory-bound sums, L2L and L2P
  // run beside P2P instead of after it
  c->p2p_started = false;
  dbuf_ensure(c->lexp, (size_t)nL * KP * esz + 16);
  fill_zero(s, c->lexp.p, esz * (size_t)nL * KP);
  run_m2l(c, KP, f32, ord, p);
  if (!c->p2p_started) launch_p2p(c, 1);
  mark(c, 13);
  // ---- L2L, coarsest level first (parents of levels < lc in FP64 in mixed precision) ----
  if (lc > 0) {
    double *L64 = dget<double>(c->lexp64, (size_t)nL * KP);
    widen(c, (const float *)c->lexp.p, (int64_t)nL * KP, L64);
    for (int d = 0; d < lc; ++d) run_trans64(c, true, KP, l2l_cells, &dp->reads[32 + d], l2l_lev[d], L64);
    narrow(c, L64, (int64_t)nL * KP, (float *)c->lexp.p);
  }
  for (int d = lc; d < D; ++d) run_trans(c, true, KP, f32, l2l_cells, &dp->reads[32 + d], l2l_lev[d]);
  mark(c, 4);
  // ---- L2P ----
  if (f32)
    hipLaunchKernelGGL((pot ? k_l2p<float, true> : k_l2p<float, false>), dim3(nblk(P.n)), dim3(kT), 0, s, p, KP, P.n,
                       (const double4 *)c->pos.p, (const int32_t *)c->pcell.p, (const int32_t *)c->cparent.p, ckey,
                       (const int32_t *)eslot, (const DevPlan *)dp, dC, (const float *)c->lexp.p, grad, pfar, ord);
  else
    hipLaunchKernelGGL((pot ? k_l2p<double, true> : k_l2p<double, false>), dim3(nblk(P.n)), dim3(kT), 0, s, p, KP,
                       P.n, (const double4 *)c->pos.p, (const int32_t *)c->pcell.p, (const int32_t *)c->cparent.p,
                       ckey, (const int32_t *)eslot, (const DevPlan *)dp, dC, (const double *)c->lexp.p, grad, pfar,
                       ord);
  NB_LAUNCH();
  mark(c, 5);
}

} // namespace nbfmm
