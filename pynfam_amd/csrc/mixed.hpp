// Split-BF16 grid stage of the mixed-precision path (device side).
//
// The grid stage of calc_hamiltonian (pnfam_hamiltonian_blas.f90:124-1262) is
// two contractions between the block elements k = (a, b) of a block matrix and
// the grid points r:
//   densities   D_c^cls(r, n) = sum_{k in cls} psi_x(r,a) psi_y(r,b) rho_k(n)
//   projection  h_k(n)        = sum_{c, r} psi_z(r,a) psi_z'(r,b) F_c^cls(k)(r, n)
// with c = (x, y) or (z, z') a pair of wavefunction tables, cls = (s_a, s_b) the
// spin class of the element and n a column (frequency, sector, re/im).  The
// products psi psi are omega-independent, so they are built once per operator
// as resident tables ("pair tables", a few GB in HBM) and both contractions
// become plain GEMMs whose large operand is streamed from HBM once per pass
// and shared by every frequency of the batch.  Operands are split into BF16
// hi + lo parts; three MFMA products (hi*hi + hi*lo + lo*hi) accumulate in FP32,
// which keeps about 16 significant bits of every product.
//
// Element order k: class-major (cls = 2 s_a + s_b), blocks in row order inside a
// class, column-major inside a block; each class range is padded to a multiple
// of MIX_KALIGN with zero table entries.

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int MIX_KALIGN = 128;   // class ranges padded to this many elements
constexpr int MIX_RALIGN = 32;    // grid points padded to this
constexpr int GT_ROWS = 128;      // table rows per workgroup
constexpr int GS_ROWS = 256;      // small-operand rows (columns n) per workgroup
constexpr int GKS = 32;           // K per LDS stage (one MFMA K step)
constexpr int GLDP = 40;          // LDS row pitch in bf16 (80 bytes)

// Projection tables: the mean fields mf(z, z', s_a, s_b) of
// pnfam_hamiltonian_blas.f90:774-1093 are non-zero for 18 pairs (z, z') (z or
// z' = 4 only with 0), and within a spin class several are equal or differ by
// a sign, so sum_{z,z'} psi_z(a) psi_z'(b) mf needs 15 (same-spin classes) or
// 16 (opposite-spin classes) table combinations, each with a combined field:
//   c 0..6: (0,0) (0,1) (1,0) (0,2) (2,0) (0,3) (3,0)      fields mf of the pair
//   c 7:    (0,4) + (4,0)                                  mf(0,4)
//   c 8:    (1,1) + (2,2) + (3,3)                          same spin mf(1,1), else mf(3,3)
//   c 9:    same spin (3,3), else (1,1)                    mf(3,3) - mf(1,1), mf(1,1) - mf(3,3)
//   c 10:   (1,2) + (2,1), else (1,2) - (2,1)              mf(1,2)
//   c 11..14: (1,3) (3,1) (2,3) (3,2)                      mf of the pair
//   c 15:   none, else (2,2)                               -, mf(2,2) - mf(3,3)
constexpr int NPROJ_C = 16;
__device__ inline bool spin_flip_class(int cls) { return cls == 1 || cls == 2; }
__device__ inline double proj_table(int cls, int c, const double (&pa)[5], const double (&pb)[5]) {
  const bool flip = spin_flip_class(cls);
  switch (c) {
    case 0: return pa[0] * pb[0];
    case 1: return pa[0] * pb[1];
    case 2: return pa[1] * pb[0];
    case 3: return pa[0] * pb[2];
    case 4: return pa[2] * pb[0];
    case 5: return pa[0] * pb[3];
    case 6: return pa[3] * pb[0];
    case 7: return pa[0] * pb[4] + pa[4] * pb[0];
    case 8: return pa[1] * pb[1] + pa[2] * pb[2] + pa[3] * pb[3];
    case 9: return flip ? pa[1] * pb[1] : pa[3] * pb[3];
    case 10: return flip ? pa[1] * pb[2] - pa[2] * pb[1] : pa[1] * pb[2] + pa[2] * pb[1];
    case 11: return pa[1] * pb[3];
    case 12: return pa[3] * pb[1];
    case 13: return pa[2] * pb[3];
    case 14: return pa[3] * pb[2];
    default: return flip ? pa[2] * pb[2] : 0.0;
  }
}
// Field of combination c for class cls (component ri), from the fields
// [208][nghl] of one sector at grid point r.
__device__ inline double proj_field(int cls, int c, const double *f, int G, int r, int ri) {
  const int sa = cls >> 1, sb = cls & 1;
  const bool flip = spin_flip_class(cls);
  auto mf = [&](int z, int zp) { return f[(size_t)(2 * mf_index(z, zp, sa, sb) + ri) * G + r]; };
  switch (c) {
    case 0: return mf(0, 0);
    case 1: return mf(0, 1);
    case 2: return mf(1, 0);
    case 3: return mf(0, 2);
    case 4: return mf(2, 0);
    case 5: return mf(0, 3);
    case 6: return mf(3, 0);
    case 7: return mf(0, 4);
    case 8: return flip ? mf(3, 3) : mf(1, 1);
    case 9: return flip ? mf(1, 1) - mf(3, 3) : mf(3, 3) - mf(1, 1);
    case 10: return mf(1, 2);
    case 11: return mf(1, 3);
    case 12: return mf(3, 1);
    case 13: return mf(2, 3);
    case 14: return mf(3, 2);
    default: return flip ? mf(2, 2) - mf(3, 3) : 0.0;
  }
}
// combinations used per class (K extent of the projection GEMM in tables)
__host__ __device__ constexpr int proj_ncomb(int cls) { return (cls == 1 || cls == 2) ? 16 : 15;}

// Density tables: per spin class the 28 densities need only nine linear
// combinations of the products psi_x(r,a) psi_y(r,b) (x, y = psi, d_r, d_phi,
// d_z psi; the products of pnfam_hamiltonian_blas.f90:255-581), because the spin
// signs are constant over a class.  kDensW[cls][c][4 x + y]: classes 0 (up,up)
// and 3 (down,down) use the same-spin combinations P, classes 1 (up,down) and 2
// (down,up) the opposite-spin combinations Q with e = +1 / -1 baked in:
//   P: 00, 02+20, 11+22+33, 01-10, 03-30, 03+30, 13+31, 23-32, 33
//   Q: 00, 02+20, e(02-20)+01+10, 11+22+33, 01-10, 03-30, 2*11+e(12-21),
//      2e*22+21-12, 31+13+e(32-23)
constexpr int NDENS_C = 9;
__constant__ int8_t kDensW[4][NDENS_C][16] = {
    // class 0 (up, up): P
    {{1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1},
     {0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0},
     {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0},
     {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0},
     {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0},
     {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1}},
    // class 1 (up, down): Q with e = +1
    {{1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
     {0, 1, 1, 0, 1, 0, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1},
     {0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0},
     {0, 0, 0, 0, 0, 2, 1, 0, 0, -1, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0, -1, 0, 0, 1, 2, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, -1, 0, 1, 1, 0}},
    // class 2 (down, up): Q with e = -1
    {{1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
     {0, 1, -1, 0, 1, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1},
     {0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0},
     {0, 0, 0, 0, 0, 2, -1, 0, 0, 1, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0, -1, 0, 0, 1, -2, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 1, -1, 0}},
    // class 3 (down, down): P
    {{1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1},
     {0, 1, 0, 0, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
     {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 0, 0, 0},
     {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0},
     {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0},
     {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, -1, 0},
     {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1}}};

__device__ inline void split_bf16(float v, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

// ---- pair tables --------------------------------------------------------
struct TabArgs {
  const double *psi[5];       // [state][r]
  const int *sa, *sb;         // [Kp] states of the element, -1 for padding
  int kbeg[5];                // class ranges of the element order
  int Kp, nghl, RP;
  int nc;                     // combos
  int dens;                   // 1: density layout [c][r][k], 0: projection layout [c][k][r]
  int single;                 // 1: the (0,0) combo only
  __bf16 *hi, *lo;
};

// One thread per (k, r) pair of one combo (blockIdx.y).  Projection layout is
// written with r fastest, density layout with k fastest.
__global__ void k_mix_tables(TabArgs T) {
  const int c = blockIdx.y;
  int x, y;
  if (T.single) { x = 0; y = 0; }
  else { x = -1; y = -1; }                   // combinations kDensW / proj_table
  const size_t n = (size_t)T.Kp * T.RP;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
    int k, r;
    if (T.dens) { k = (int)(e % T.Kp); r = (int)(e / T.Kp); }
    else { r = (int)(e % T.RP); k = (int)(e / T.RP); }
    const int a = T.sa[k], b = T.sb[k];
    float v = 0.0f;
    if (a >= 0 && r < T.nghl) {
      if (x >= 0) {
        v = (float)(T.psi[x][(size_t)a * T.nghl + r] * T.psi[y][(size_t)b * T.nghl + r]);
      } else {
        const int cls = k < T.kbeg[1] ? 0 : k < T.kbeg[2] ? 1 : k < T.kbeg[3] ? 2 : 3;
        double pa[5], pb[5];
        for (int q = 0; q < 5; ++q) {
          pa[q] = T.psi[q][(size_t)a * T.nghl + r];
          pb[q] = T.psi[q][(size_t)b * T.nghl + r];
        }
        double acc = 0.0;
        if (T.dens) {
          for (int xy = 0; xy < 16; ++xy) acc += (double)kDensW[cls][c][xy] * pa[xy >> 2] * pb[xy & 3];
        } else {
          acc = proj_table(cls, c, pa, pb);
        }
        v = (float)acc;
      }
    }
    __bf16 h, l;
    split_bf16(v, h, l);
    // density layout [c][r][k]; projection layout [k][c][r] (K = (c, r) contiguous per row k)
    const size_t o = T.dens ? (size_t)c * n + e : ((size_t)k * T.nc + c) * T.RP + r;
    T.hi[o] = h;
    T.lo[o] = l;
  }
}

// FP32 fill (output planes of classes without elements)
__global__ void k_fill_f32(float *__restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

// Projection results O[k][n] (FP32, n fastest) into the FP64 block matrices:
// column n = (slot * nsec + sec) * 2 + ri, element elem[k] of quadrant q(sec).
// A 64 (k) x 64 (n) tile goes through LDS: the reads run along n, the writes
// along k (consecutive elements of a block column).
__global__ void __launch_bounds__(256) k_mix_scatter(const float *__restrict__ O, long ldO,
                                                     const int *__restrict__ elem, int Kp, int nS, int nsec, int q0,
                                                     int q1, const int *__restrict__ act, double *__restrict__ hsp,
                                                     long nxy) {
  __shared__ float tile[64][65];
  const int k0 = blockIdx.x * 64, s0 = blockIdx.y * 64;
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int kk = idx >> 6, ss = idx & 63;
    const int k = k0 + kk, s = s0 + ss;
    tile[kk][ss] = (k < Kp && s < nS) ? O[(size_t)k * ldO + s] : 0.0f;
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < 64 * 64; idx += 256) {
    const int ss = idx >> 6, kk = idx & 63;
    const int k = k0 + kk, s = s0 + ss;
    if (k >= Kp || s >= nS) continue;
    const int e = elem[k];
    if (e < 0) continue;
    const int ri = s & 1;
    const int sec = nsec == 2 ? (s >> 1) & 1 : 0;
    const int slot = s / (2 * nsec);
    const int q = sec ? q1 : q0;
    hsp[((size_t)act[slot] * 8 + 2 * q + ri) * nxy + e] = (double)tile[kk][ss];
  }
}

// ---- operand conversion -------------------------------------------------
// Bd[n][k] (k fastest) from the FP64 block matrices in the reference layout.
// n = (slot * nsec + sec) * 2 + ri; sector sec reads quadrant q_of_sec[sec].
struct CvtRhoArgs {
  const double *rsp;          // [w][8][nxy]
  const int *elem;            // [Kp]
  const int *act;
  long nxy;
  int Kp, nsec, q0, q1, ncap;
  __bf16 *hi, *lo;            // [ncap][Kp]
};

__global__ void k_mix_rho(CvtRhoArgs A) {
  const int slot = blockIdx.y;
  const int w = A.act[slot];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < A.Kp; k += gridDim.x * blockDim.x) {
    const int e = A.elem[k];
    for (int sec = 0; sec < A.nsec; ++sec) {
      const int q = sec ? A.q1 : A.q0;
      for (int ri = 0; ri < 2; ++ri) {
        const float v = e >= 0 ? (float)A.rsp[((size_t)w * 8 + 2 * q + ri) * A.nxy + e] : 0.0f;
        __bf16 h, l;
        split_bf16(v, h, l);
        const size_t o = (size_t)((slot * A.nsec + sec) * 2 + ri) * A.Kp + k;
        A.hi[o] = h;
        A.lo[o] = l;
      }
    }
  }
}

// F[cls][n][c][r] (r fastest) from the FP64 fields [w][sec][208][r], times the
// factor 2 of the projection (pnfam_hamiltonian_blas.f90:1106-1160, 1240-1262).
// pair != 0: the four pairing fields of one sector, F[cls][n][r] with
// n = slot * 2 + ri.  Otherwise the 18 mean-field pairs for both sectors.
struct CvtFldArgs {
  const double *flds;
  const int *act;
  int nghl, RP, ncap;
  int pair, sec;
  __bf16 *hi, *lo;
};

__global__ void k_mix_fields(CvtFldArgs A) {
  const int slot = blockIdx.y;
  const int w = A.act[slot];
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= A.RP) return;
  const bool live = r < A.nghl;
  const size_t plane = (size_t)A.ncap * A.RP;
  if (A.pair) {
    const double *f = A.flds + (((size_t)w * 2 + A.sec) * 2 * NFLD) * A.nghl;
    for (int cls = 0; cls < 4; ++cls) {
      const int sa = cls >> 1, sb = cls & 1;
      // pairing field index of (s_a, s_b): (up,up) 0, (down,up) 1, (up,down) 2, (down,down) 3
      const int idx = 100 + (sb == 0 ? (sa == 0 ? 0 : 1) : (sa == 0 ? 2 : 3));
      for (int ri = 0; ri < 2; ++ri) {
        const float v = live ? (float)(2.0 * f[(size_t)(2 * idx + ri) * A.nghl + r]) : 0.0f;
        __bf16 h, l;
        split_bf16(v, h, l);
        const size_t o = (size_t)cls * plane + (size_t)(slot * 2 + ri) * A.RP + r;
        A.hi[o] = h;
        A.lo[o] = l;
      }
    }
    return;
  }
  for (int sec = 0; sec < 2; ++sec) {
    const double *f = A.flds + (((size_t)w * 2 + sec) * 2 * NFLD) * A.nghl;
    // unrolled: the combination and class are compile-time, so proj_field
    // reduces to its loads (a runtime switch kept the fields in scratch)
#pragma unroll
    for (int c = 0; c < NPROJ_C; ++c) {
#pragma unroll
      for (int cls = 0; cls < 4; ++cls) {
#pragma unroll
        for (int ri = 0; ri < 2; ++ri) {
          const float v = live ? (float)(2.0 * proj_field(cls, c, f, A.nghl, r, ri)) : 0.0f;
          __bf16 h, l;
          split_bf16(v, h, l);
          const size_t o = (((size_t)cls * A.ncap + (slot * 2 + sec) * 2 + ri) * NPROJ_C + c) * A.RP + r;
          A.hi[o] = h;
          A.lo[o] = l;
        }
      }
    }
  }
}

// ---- the split GEMM -------------------------------------------------------
// O[t][s] = sum_{seg, kk} T[t][seg, kk] * S[s][seg, kk] with both operands
// K-contiguous BF16 hi/lo.  One job per spin class.
struct GemmJob {
  int t0, tcount;     // table rows of this job
  int k0, klen;       // K range inside each segment (klen multiple of GKS)
  int cls, pad;       // spin class: selects the small operand and output planes
};

struct GemmArgs {
  const __bf16 *Th, *Tl;
  long ldT, segT;
  const __bf16 *Sh, *Sl;
  long ldS, segS;
  int nseg, nS, nst;
  long s_cls, o_cls;  // per-class offsets of the small operand and of the output
  const GemmJob *jobs;
  int mode;           // 0: FP32 O[t][s]; 1: FP64 scatter into the block matrices
  float *O;
  long ldO;
  double *hsp;        // mode 1: [w][8][nxy]
  const int *elem;    // mode 1: [Kp]
  const int *act;
  long nxy;
  int nsec, q0, q1;
};

// Workgroup: 8 waves, 128 table rows x 256 columns; each wave 64 x 64 (4 x 4
// MFMA tiles of v_mfma_f32_16x16x32_bf16).  The flat block index is dealt to
// the 8 XCDs round-robin; it is decoded so that all column tiles of one table
// tile run on one XCD and share that table tile through its L2.
__global__ void __launch_bounds__(512) k_mix_gemm(GemmArgs G) {
  __shared__ __attribute__((aligned(16))) __bf16 sT[2][GT_ROWS][GLDP];
  __shared__ __attribute__((aligned(16))) __bf16 sS[2][GS_ROWS][GLDP];
  const GemmJob J = G.jobs[blockIdx.y];
  const int b = blockIdx.x;
  const int xcd = b & 7, loc = b >> 3;
  const int st = loc % G.nst;
  const int tt = (loc / G.nst) * 8 + xcd;
  if (tt * GT_ROWS >= J.tcount) return;   // whole workgroup leaves before any barrier
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wt = wave & 1, ws = wave >> 1;
  const int fi = lane & 15, fg = lane >> 4;
  // wave-uniform activity of the 4 column tiles of this wave
  const int s_wave0 = st * GS_ROWS + ws * 64;
  int nj = (G.nS - s_wave0 + 15) / 16;
  nj = nj < 0 ? 0 : (nj > 4 ? 4 : nj);
  int ni = (J.tcount - (tt * GT_ROWS + wt * 64) + 15) / 16;
  ni = ni < 0 ? 0 : (ni > 4 ? 4 : ni);

  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  // staging assignments: table tile 128 x 32 = 512 chunks of 8 (one per thread),
  // small tile 256 x 32 = 1024 chunks (two per thread)
  const int trow = tid >> 2, tch = tid & 3;
  const bool t_ok = tt * GT_ROWS + trow < J.tcount;
  const long t_base = (long)(J.t0 + tt * GT_ROWS + (t_ok ? trow : 0)) * G.ldT + J.k0 + tch * 8;
  const long s_off = (long)J.cls * G.s_cls;
  const long o_off = (long)J.cls * G.o_cls;
  int srow[2], sch[2];
  bool s_ok[2];
  long s_base[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int cidx = tid + q * 512;
    srow[q] = cidx >> 2; sch[q] = cidx & 3;
    const int s = st * GS_ROWS + srow[q];
    s_ok[q] = s < G.nS;
    s_base[q] = s_off + (long)(s_ok[q] ? s : 0) * G.ldS + J.k0 + sch[q] * 8;
  }
  const v8bf z8 = {0, 0, 0, 0, 0, 0, 0, 0};
  v8bf rth, rtl, rsh[2], rsl[2];
  auto load = [&](int sg, int kk) {
    const long to = t_base + (long)sg * G.segT + kk;
    rth = t_ok ? *(const v8bf *)(G.Th + to) : z8;
    rtl = t_ok ? *(const v8bf *)(G.Tl + to) : z8;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const long so = s_base[q] + (long)sg * G.segS + kk;
      rsh[q] = s_ok[q] ? *(const v8bf *)(G.Sh + so) : z8;
      rsl[q] = s_ok[q] ? *(const v8bf *)(G.Sl + so) : z8;
    }
  };
  auto store = [&]() {
    *(v8bf *)&sT[0][trow][tch * 8] = rth;
    *(v8bf *)&sT[1][trow][tch * 8] = rtl;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      *(v8bf *)&sS[0][srow[q]][sch[q] * 8] = rsh[q];
      *(v8bf *)&sS[1][srow[q]][sch[q] * 8] = rsl[q];
    }
  };
  const int nk = J.klen / GKS;
  const int nsteps = G.nseg * nk;
  if (nsteps > 0) {
    load(0, 0);
    store();
  }
  __syncthreads();
  for (int step = 0; step < nsteps; ++step) {
    const bool more = step + 1 < nsteps;
    if (more) load((step + 1) / nk, ((step + 1) % nk) * GKS);   // in flight during the MFMAs
    if (ni > 0 && nj > 0) {
      v8bf ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        ah[i] = *(const v8bf *)&sT[0][wt * 64 + i * 16 + fi][fg * 8];
        al[i] = *(const v8bf *)&sT[1][wt * 64 + i * 16 + fi][fg * 8];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        bh[j] = *(const v8bf *)&sS[0][ws * 64 + j * 16 + fi][fg * 8];
        bl[j] = *(const v8bf *)&sS[1][ws * 64 + j * 16 + fi][fg * 8];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i >= ni) break;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (j >= nj) break;
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
    if (more) {
      store();
      __syncthreads();
    }
  }
  // epilogue: lane holds D[t = 4 (lane / 16) + v][s = lane % 16] of every tile
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= ni) break;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j >= nj) break;
      const int s = s_wave0 + j * 16 + (lane & 15);
      if (s >= G.nS) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int tl = tt * GT_ROWS + wt * 64 + i * 16 + 4 * (lane >> 4) + v;
        if (tl >= J.tcount) continue;
        if (G.mode == 0) {
          G.O[o_off + (long)tl * G.ldO + s] = acc[i][j][v];
        } else {
          const int e = G.elem[J.t0 + tl];
          if (e < 0) continue;
          const int ri = s & 1;
          const int sec = G.nsec == 2 ? (s >> 1) & 1 : 0;
          const int slot = s / (2 * G.nsec);
          const int q = sec ? G.q1 : G.q0;
          G.hsp[((size_t)G.act[slot] * 8 + 2 * q + ri) * G.nxy + e] = (double)acc[i][j][v];
        }
      }
    }
  }
}

// ---- densities from the class-summed contractions -------------------------
// Of[cls][c][r][n] (mean-field pairs c = 4x + y, n = (slot*2 + sec)*2 + ri) and
// Ok{0,1}[cls][r][n'] (pairing, n' = slot*2 + ri) -> dens[w][sec][56][r],
// with the products of pnfam_hamiltonian_blas.f90:255-581 per spin of b.
struct CombArgs {
  const float *Of, *Ok0, *Ok1;
  long ldf, ldk;             // row pitch (columns) of Of and Ok
  int RP, nghl, na;
  const double *wdcori;
  double *dens;              // [w][2][56][nghl]
  const int *act;
};

// One thread per (grid point, frequency slot, sector) of a 16 x 16 tile; the
// sixteen threads of a row run along (slot, sector), the contiguous direction
// of Of and Ok, and the results go out through LDS so that sixteen threads
// write sixteen consecutive grid points of one density.
__global__ void __launch_bounds__(256) k_mix_dens(CombArgs A) {
  __shared__ double sT[16][17];
  const int tj = threadIdx.x & 15, tr = threadIdx.x >> 4;
  const int r0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  const int r = min(r0 + tr, A.nghl - 1);            // clamped: out-of-range rows are not written
  const int j = min(j0 + tj, 2 * A.na - 1);
  const int slot = j >> 1, sec = j & 1;
  double acc[2 * NDENS];
#pragma unroll
  for (int d = 0; d < 2 * NDENS; ++d) acc[d] = 0.0;
  const size_t plane_f = (size_t)A.RP * A.ldf;
  const size_t plane_k = (size_t)A.RP * A.ldk;
  const float *Ok = sec ? A.Ok1 : A.Ok0;
  for (int sb = 0; sb < 2; ++sb) {
    const double s = sb == 0 ? 1.0 : -1.0, e = -s;
    const int cd = 2 * sb + sb, co = 2 * (1 - sb) + sb;   // (s_a = s_b, s_b), (s_a != s_b, s_b)
    double Pr9[NDENS_C], Pi9[NDENS_C], Qr9[NDENS_C], Qi9[NDENS_C];
#pragma unroll
    for (int c = 0; c < NDENS_C; ++c) {
      const float *pd = A.Of + (size_t)(cd * NDENS_C + c) * plane_f + (size_t)r * A.ldf + 2 * j;
      const float *po = A.Of + (size_t)(co * NDENS_C + c) * plane_f + (size_t)r * A.ldf + 2 * j;
      Pr9[c] = pd[0]; Pi9[c] = pd[1];
      Qr9[c] = po[0]; Qi9[c] = po[1];
    }
#define ACC(d, ri) acc[2 * (d) + (ri)]
    // same spin (P combinations); i*z adds (-Im z, Re z)
    ACC(D_RHO, 0) += Pr9[0]; ACC(D_RHO, 1) += Pi9[0];
    ACC(D_SZ, 0) += s * Pr9[0]; ACC(D_SZ, 1) += s * Pi9[0];
    ACC(D_JP, 0) += 0.5 * Pr9[1]; ACC(D_JP, 1) += 0.5 * Pi9[1];
    ACC(D_TJPZ, 0) += 0.5 * s * Pr9[1]; ACC(D_TJPZ, 1) += 0.5 * s * Pi9[1];
    ACC(D_TAU, 0) += Pr9[2]; ACC(D_TAU, 1) += Pi9[2];
    ACC(D_TZ, 0) += s * Pr9[2]; ACC(D_TZ, 1) += s * Pi9[2];
    ACC(D_JR, 0) -= 0.5 * Pi9[3]; ACC(D_JR, 1) += 0.5 * Pr9[3];
    ACC(D_TJRZ, 0) -= 0.5 * s * Pi9[3]; ACC(D_TJRZ, 1) += 0.5 * s * Pr9[3];
    ACC(D_JZ, 0) -= 0.5 * Pi9[4]; ACC(D_JZ, 1) += 0.5 * Pr9[4];
    ACC(D_TJZZ, 0) -= 0.5 * s * Pi9[4]; ACC(D_TJZZ, 1) += 0.5 * s * Pr9[4];
    ACC(D_GS, 0) += s * Pr9[5]; ACC(D_GS, 1) += s * Pi9[5];
    ACC(D_FR, 0) += 0.5 * s * Pr9[6]; ACC(D_FR, 1) += 0.5 * s * Pi9[6];
    ACC(D_FP, 0) -= 0.5 * s * Pi9[7]; ACC(D_FP, 1) += 0.5 * s * Pr9[7];
    ACC(D_FZ, 0) += s * Pr9[8]; ACC(D_FZ, 1) += s * Pi9[8];
    // opposite spin (Q combinations, e = -s)
    ACC(D_SR, 0) += Qr9[0]; ACC(D_SR, 1) += Qi9[0];
    ACC(D_SP, 0) -= e * Qi9[0]; ACC(D_SP, 1) += e * Qr9[0];
    ACC(D_TJPR, 0) += 0.5 * Qr9[1]; ACC(D_TJPR, 1) += 0.5 * Qi9[1];
    ACC(D_TJPP, 0) -= 0.5 * e * Qi9[1]; ACC(D_TJPP, 1) += 0.5 * e * Qr9[1];
    ACC(D_GS, 0) += Qr9[2]; ACC(D_GS, 1) += Qi9[2];
    ACC(D_TR, 0) += Qr9[3]; ACC(D_TR, 1) += Qi9[3];
    ACC(D_TP, 0) -= e * Qi9[3]; ACC(D_TP, 1) += e * Qr9[3];
    ACC(D_TJRR, 0) -= 0.5 * Qi9[4]; ACC(D_TJRR, 1) += 0.5 * Qr9[4];
    ACC(D_TJRP, 0) -= 0.5 * e * Qr9[4]; ACC(D_TJRP, 1) -= 0.5 * e * Qi9[4];
    ACC(D_TJZR, 0) -= 0.5 * Qi9[5]; ACC(D_TJZR, 1) += 0.5 * Qr9[5];
    ACC(D_TJZP, 0) -= 0.5 * e * Qr9[5]; ACC(D_TJZP, 1) -= 0.5 * e * Qi9[5];
    ACC(D_FR, 0) += 0.5 * Qr9[6]; ACC(D_FR, 1) += 0.5 * Qi9[6];
    ACC(D_FP, 0) -= 0.5 * Qi9[7]; ACC(D_FP, 1) += 0.5 * Qr9[7];
    ACC(D_FZ, 0) += 0.5 * Qr9[8]; ACC(D_FZ, 1) += 0.5 * Qi9[8];
#undef ACC
    // pairing densities (pnfam_hamiltonian_blas.f90:584-640)
    const float *kd = Ok + (size_t)cd * plane_k + (size_t)r * A.ldk + 2 * slot;
    const float *ko = Ok + (size_t)co * plane_k + (size_t)r * A.ldk + 2 * slot;
    double pr = 2.0 * ko[0], pi = 2.0 * ko[1];
    acc[2 * D_RB] -= s * pr; acc[2 * D_RB + 1] -= s * pi;
    acc[2 * D_SBZ] += pr; acc[2 * D_SBZ + 1] += pi;
    pr = 2.0 * kd[0]; pi = 2.0 * kd[1];
    acc[2 * D_SBR] += s < 0 ? pr : -pr; acc[2 * D_SBR + 1] += s < 0 ? pi : -pi;
    acc[2 * D_SBP] += pi; acc[2 * D_SBP + 1] -= pr;
  }
  const double wc = A.wdcori[r];
  // writer view of the tile: wj = (slot, sector) row, wr = grid point
  const int wj = threadIdx.x >> 4, wr = threadIdx.x & 15;
  const int jw = j0 + wj, rw = r0 + wr;
  const bool wok = jw < 2 * A.na && rw < A.nghl;
  double *out = wok ? A.dens + (((size_t)A.act[jw >> 1] * 2 + (jw & 1)) * 2 * NDENS) * A.nghl + rw : nullptr;
#pragma unroll
  for (int d = 0; d < 2 * NDENS; ++d) {
    sT[tj][tr] = acc[d] * wc;
    __syncthreads();
    if (wok) out[(size_t)d * A.nghl] = sT[wj][wr];
    __syncthreads();
  }
}
