// Chunked line block ILU(0) of A = I/dt - J on the 13-point stencil.
//
// The i-range is cut into chunks of W columns.  Inside a chunk the cells are
// ordered column by column with j fastest (o = (i - i0) njc + j), so every
// wall-normal line is contiguous and its banded part is factored exactly;
// couplings between chunks are dropped (block Jacobi over chunks).  In that
// order the lower neighbours of (i, j) are (-2,0), (-1,-1), (-1,0), (-1,+1),
// (0,-2), (0,-1) and all of them lie on an earlier level L = j + 2 (i - i0);
// a chunk has njc + 2W - 2 levels of at most W cells.  One work-group per
// chunk walks the levels with a barrier between them.
//
// Factors are stored in FP32 (the preconditioner only has to approximate
// A^-1; FGMRES runs in FP64), per chunk in level order, one record per level:
// rec[(((chunk * nlev + L) * NSLOT + s) * NV + r) * W + c) * NV + m] for entry
// (r, m) of the block in slot s of the cell in column c of the chunk (row
// j = L - 2c): slot 0 the inverse of the factored diagonal block, lower slots
// L_rn (unit lower factor), upper slots U_ru.  A block row is NV contiguous
// floats and the rows r of a level's cells sit next to each other, so the
// thread for (r, c) reads its row with two wide loads and a wave's loads are
// coalesced.  A second copy holds every block transposed, for M^-T: a column
// of a block is then read like a row, with the same loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "sg_linalg.hpp"

namespace sg {

// Work-group barrier that orders LDS only: the level loop keeps the next
// level's global loads in flight across it (the usual barrier also waits for
// every outstanding global load).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Compile-time stencil tables for the apply kernel (fully unrolled loops
// fold them away): offsets, the pass's neighbour slots, and the slot of the
// opposite offset.
namespace ct {
constexpr int Di[NSLOT] = {0, 0, 0, -1, 1, -1, 1, -1, 1, 0, 0, -2, 2};
constexpr int Dj[NSLOT] = {0, -1, 1, 0, 0, -1, -1, 1, 1, -2, 2, 0, 0};
constexpr int Low[6] = {11, 5, 3, 7, 9, 1};
constexpr int Up[6] = {2, 10, 6, 4, 8, 12};
constexpr int Back[NSLOT] = {0, 2, 1, 4, 3, 8, 7, 6, 5, 10, 9, 12, 11};
// The 3x3-box pattern (slots 0-8; the +-2 slots dropped): four lower and
// four upper neighbours, in the same order as above.
constexpr int Low9[4] = {5, 3, 7, 1};
constexpr int Up9[4] = {2, 6, 4, 8};
template <bool NINE>
constexpr int nn() { return NINE ? 4 : 6; }
template <bool NINE>
constexpr int low(int n) { return NINE ? Low9[n] : Low[n]; }
template <bool NINE>
constexpr int up(int n) { return NINE ? Up9[n] : Up[n]; }
constexpr bool in_box(int s) { return s >= 0 && s < 9; }
constexpr int slot(int di, int dj) {
  for (int s = 0; s < NSLOT; ++s)
    if (Di[s] == di && Dj[s] == dj) return s;
  return -1;
}
}  // namespace ct

// Factorisation, one level after another.  Thread (c, e) holds entry e of
// every stencil block of its cell's row in registers.  A row's lower
// neighbours lie on the four previous levels (L-1 .. L-4), whose factored
// diagonal inverses and upper blocks are kept in an LDS ring, so the level
// loop reads the 13 blocks of its own cells (gathered in level order by
// cilu_gather_kernel) from memory and nothing else; the elimination exchanges
// one 5x5 block per cell through LDS per lower neighbour.
// Record size of one level, in floats.
template <int NV, int W>
constexpr int cilu_rs() { return NSLOT * NV * NV * W; }

// Gathers the factorisation's operands into level order: for chunk, level L,
// slot s, block entry e and column c, entry e of the block of I/dt - J in row
// (i0 + c, L - 2c) and neighbour slot s, or zero where the cell or the
// neighbour lies outside the chunk or the grid.  Consecutive columns are
// contiguous, so the factorisation reads a level's operands in runs of W
// doubles.  Read in place, a level's W cells lie on a diagonal of rows 2 ldi
// apart: every 8-byte read took its own line, used again only ~W levels later
// after L2 had evicted it.  Here all levels are gathered at once and the
// levels that share a line are read together.
template <int NV, int W>
__global__ __launch_bounds__(256) void cilu_gather_kernel(int nic, int njc, int ldi, int nlev, const double* idt,
                                                          const double* blocks, double* ops, size_t total) {
  constexpr int B = NV * NV;
  const size_t plane = (size_t)njc * ldi;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    size_t u = t;
    const int c = (int)(u % W);
    u /= W;
    const int e = (int)(u % B);
    u /= B;
    const int s = (int)(u % NSLOT);
    u /= NSLOT;
    const int L = (int)(u % (size_t)nlev), chunk = (int)(u / (size_t)nlev);
    const int i0 = chunk * W, wk = min(W, nic - i0), j = L - 2 * c;
    const int cn = c + ct::Di[s], jn = j + ct::Dj[s];
    double v = 0.0;
    if (c < wk && j >= 0 && j < njc && cn >= 0 && cn < wk && jn >= 0 && jn < njc) {
      const size_t cell = (size_t)j * ldi + i0 + c;
      v = -blocks[(size_t)(s * B + e) * plane + cell];
      if (s == 0 && e / NV == e % NV) v += idt[cell];
    }
    ops[t] = v;
  }
}

// floats of one record copy (plain or transposed)
static size_t cilu_copy_floats(int nic, int njc, int nv, int w) {
  const int nchunk = (nic + w - 1) / w, nlev = njc + 2 * w - 2;
  return (size_t)nchunk * nlev * NSLOT * nv * nv * w;
}

// NINE: ILU(0) on the 3x3-box pattern (the +-2 blocks of A and their fill
// dropped; with the first-order Jacobian of lhs_order 1 those slots hold
// fill only), else on the 13-point stencil.
template <int NV, int W, bool NINE>
__global__ __launch_bounds__(W * NV * NV) void cilu_factor_kernel(int nic, int njc, int nlev, const double* ops,
                                                                  float* rec, size_t tcopy) {
  constexpr int B = NV * NV, RS = cilu_rs<NV, W>(), NR = 4, NN = ct::nn<NINE>();
  __shared__ float ring[NR][7][W][B];  // per level: diagonal inverse, then the six upper blocks (ct::Up order)
  __shared__ double Ab[W][B];
  __shared__ double Lb[W][B];
  __shared__ double cb[2][W][NV];
  const int chunk = blockIdx.x, i0 = chunk * W, wk = min(W, nic - i0);
  const int t = threadIdx.x, c = t / B, e = t % B, r = e / NV, q = e % NV;
  float* R = rec + (size_t)chunk * nlev * RS;
  float* RT = R + tcopy;
  auto inside = [&](int cc, int jj) { return cc >= 0 && cc < wk && jj >= 0 && jj < njc; };
  for (int L = 0; L < nlev; ++L) {
    const int j = L - 2 * c;
    const bool valid = inside(c, j);
    // this row's 13 blocks of I/dt - J, gathered in level order (zero where
    // the cell or the neighbour is outside)
    const double* opl = ops + ((size_t)(chunk * nlev + L) * NSLOT * B + e) * W + c;
    double a[NSLOT];
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) a[s] = opl[(size_t)s * B * W];
#pragma unroll
    for (int qi = 0; qi < NN; ++qi) {
      const int s = ct::low<NINE>(qi), di = ct::Di[s], dj = ct::Dj[s];
      const int cn = c + di, jn = j + dj;
      const bool nok = valid && inside(cn, jn);
      const int ln = (L + dj + 2 * di) & (NR - 1);  // the neighbour's level in the ring
      const int cs = nok ? cn : c;
      Ab[c][e] = a[s];
      lds_barrier();
      double l = 0.0;
      if (nok) {
#pragma unroll
        for (int k = 0; k < NV; ++k) l += Ab[c][r * NV + k] * (double)ring[ln][0][cs][k * NV + q];
      }
      a[s] = l;
      Lb[c][e] = l;
      lds_barrier();
      if (nok) {
#pragma unroll
        for (int ui = 0; ui < NN; ++ui) {
          const int ts = ct::up<NINE>(ui);
          const int ri = di + ct::Di[ts], rj = dj + ct::Dj[ts];
          const int su = ct::slot(ri, rj);
          if (su < 0 || (NINE && !ct::in_box(su)) || !inside(c + ri, j + rj)) continue;
          double sum = 0.0;
#pragma unroll
          for (int k = 0; k < NV; ++k) sum += Lb[c][r * NV + k] * (double)ring[ln][1 + ui][cs][k * NV + q];
          a[su] -= sum;
        }
      }
    }
    // the level's L and U blocks are final: store them (record, transposed
    // record and ring)
    float* RL = R + (size_t)L * RS;
    float* RTL = RT + (size_t)L * RS;
    const int lr = L & (NR - 1);
    // one per-thread offset for every slot and a uniform slot stride (offsets
    // computed per slot were kept live across the level loop and spilled)
    constexpr int SS = NV * W * NV;
    const unsigned on = (unsigned)((r * W + c) * NV + q), ot = (unsigned)((q * W + c) * NV + r);
    auto at_n = [&](int s) { return (size_t)s * SS + on; };
    auto at_t = [&](int s) { return (size_t)s * SS + ot; };
#pragma unroll
    for (int s = 1; s < (NINE ? 9 : NSLOT); ++s) {  // the box pattern never reads the +-2 slots
      const float v = valid ? (float)a[s] : 0.0f;
      RL[at_n(s)] = v;
      RTL[at_t(s)] = v;
    }
#pragma unroll
    for (int ui = 0; ui < NN; ++ui) ring[lr][1 + ui][c][e] = valid ? (float)a[ct::up<NINE>(ui)] : 0.0f;
    // inverse of the factored diagonal block: Gauss-Jordan with partial
    // pivoting on [D | I], column-parallel (thread e < 2 NV of a cell owns
    // column e; the pivot column is broadcast through LDS each step)
    Ab[c][e] = a[0];
    lds_barrier();
    double col[NV];
#pragma unroll
    for (int rr = 0; rr < NV; ++rr) col[rr] = (e < NV) ? Ab[c][rr * NV + e] : ((e - NV == rr) ? 1.0 : 0.0);
#pragma unroll
    for (int p = 0; p < NV; ++p) {
      if (e == p) {
#pragma unroll
        for (int rr = 0; rr < NV; ++rr) cb[p & 1][c][rr] = col[rr];
      }
      lds_barrier();
      double cp[NV];
#pragma unroll
      for (int rr = 0; rr < NV; ++rr) cp[rr] = cb[p & 1][c][rr];
      int piv = p;
      double best = fabs(cp[p]);
#pragma unroll
      for (int rr = p + 1; rr < NV; ++rr) {
        const double v = fabs(cp[rr]);
        if (v > best) { best = v; piv = rr; }
      }
#pragma unroll
      for (int rr = p + 1; rr < NV; ++rr) {
        const bool sw = piv == rr;
        const double x = cp[p], y = cp[rr], u = col[p], w = col[rr];
        cp[p] = sw ? y : x; cp[rr] = sw ? x : y;
        col[p] = sw ? w : u; col[rr] = sw ? u : w;
      }
      const double ip = 1.0 / cp[p];
      col[p] *= ip;
#pragma unroll
      for (int rr = 0; rr < NV; ++rr) {
        if (rr == p) continue;
        const double f = cp[rr];
        col[rr] = (f != 0.0) ? col[rr] - f * col[p] : col[rr];
      }
    }
    if (e >= NV && e < 2 * NV) {
#pragma unroll
      for (int rr = 0; rr < NV; ++rr) Lb[c][rr * NV + e - NV] = col[rr];
    }
    lds_barrier();
    const float dinv = valid ? (float)Lb[c][e] : 0.0f;
    RL[at_n(0)] = dinv;
    RTL[at_t(0)] = dinv;
    ring[lr][0][c][e] = dinv;
  }
}

// One stored row of a block (NV contiguous floats), widened to double.  The
// row is read as one unit so that it takes two wide loads (16 + 4 bytes for
// NV = 5) rather than NV single-float loads: the level loop issues few
// enough memory instructions that their address processing does not pace it.
template <int NV>
struct FRow {
  float v[NV];
};
template <int NV, int W>
__device__ __forceinline__ void load_row(const float* p, double* out) {
  const FRow<NV> row = *reinterpret_cast<const FRow<NV>*>(p);
#pragma unroll
  for (int m = 0; m < NV; ++m) out[m] = row.v[m];
}

// One level's operands of thread (k, c): f[n*NV + m] for the six neighbours n
// of the pass, d[m] for the diagonal inverse (row k, or column k when
// transposed; only in the pass that applies it), x the right-hand side.  R is the record copy of the pass: the
// plain one for M^-1, the transposed one for M^-T, where a column of a block
// is a stored row.  The factorisation stores exact zeros for neighbours
// outside the chunk or the grid and for invalid cells (the transposed pass
// reads the cell's own slot, zero for an absent neighbour), so nothing is
// masked and all the loads of a level are in flight together.
template <int NV, int W, bool NINE, bool TR, bool ASC, bool DIAG>
__device__ __forceinline__ void cilu_fetch(const float* R, int nlev, int njc, int ldi, size_t plane, int i0, int wk,
                                           int k, int c, int L, const double* src, double* f, double* d, double& x) {
  constexpr int RS = cilu_rs<NV, W>();
  const int j = L - 2 * c;
  const bool lok = L >= 0 && L < nlev;
  const bool valid = lok && c < wk && j >= 0 && j < njc;
  const float* RL = R + (size_t)(lok ? L : 0) * RS;
  auto row = [&](const float* Rlev, int s, int r, int col) { return Rlev + ((size_t)(s * NV + r) * W + col) * NV; };
#pragma unroll
  for (int n = 0; n < ct::nn<NINE>(); ++n) {
    const int s = ASC ? ct::low<NINE>(n) : ct::up<NINE>(n);
    if (!TR) {
      load_row<NV, W>(row(RL, s, k, c), f + n * NV);
    } else {
      // column k of the neighbour's block that points back to this cell
      const int cn = c + ct::Di[s], jn = j + ct::Dj[s];
      const bool nok = valid && cn >= 0 && cn < wk && jn >= 0 && jn < njc;
      const float* p = nok ? row(R + (size_t)(jn + 2 * cn) * RS, ct::Back[s], k, cn) : row(RL, s, k, c);
      load_row<NV, W>(p, f + n * NV);
    }
  }
  if (DIAG) load_row<NV, W>(row(RL, 0, k, c), d);
  x = src[k * plane + (valid ? (size_t)j * ldi + i0 + c : 0)];
}

// One level of a triangular pass with its operands already in registers.
// Returns the thread's solved value (stored to z by the caller).
template <int NV, int W, bool NINE, bool ASC, bool DIAG>
__device__ __forceinline__ double cilu_level(int L, int njc, int wk, int k, int c, const double* f, const double* d,
                                             double x, double (*ring)[NV][W], double (*tmp)[W]) {
  constexpr int NR = 8;
  const int j = L - 2 * c;
  const bool valid = c < wk && j >= 0 && j < njc;
  double a0 = x, a1 = 0.0;
#pragma unroll
  for (int n = 0; n < ct::nn<NINE>(); ++n) {
    const int s = ASC ? ct::low<NINE>(n) : ct::up<NINE>(n);
    const int cn = c + ct::Di[s], jn = j + ct::Dj[s];
    // out-of-chunk / out-of-grid neighbours carry f = 0: any in-range ring
    // entry (all finite) will do
    const int cc = min(max(cn, 0), W - 1), jr = jn & (NR - 1);
    double t = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m) t += f[n * NV + m] * ring[jr][m][cc];
    if (n & 1) a1 += t;
    else a0 -= t;
  }
  double acc = a0 - a1;
  if (DIAG) {
    if (valid) tmp[k][c] = acc;
    lds_barrier();
    double v = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m) v += d[m] * tmp[m][c];
    acc = v;
  }
  if (valid) ring[j & (NR - 1)][k][c] = acc;
  lds_barrier();
  return acc;
}

// Operand sets in rotation: the loads of a level are issued NBUF - 1 levels
// ahead of their use.  Two sets measured best; deeper prefetch does not pay,
// because a level is paced by how fast the chunk's factor records reach one
// CU (about 22 KB per level and pass at W = 32), not by the latency of one
// load.
constexpr int NBUF = 2;

// Helper work-groups: per chunk, NHELP work-groups on the same XCD (work-groups
// are dealt to the 8 XCDs round-robin by index) load the factor lines of the
// levels ahead of the solving work-group, so that its loads hit that XCD's L2
// instead of HBM.  They only load (nothing they do affects the result), at
// most HLEAD steps ahead of the progress the solving work-group reports every
// PROG_EVERY steps, HB of their levels per batch.  Their wait for progress is
// bounded, so they always finish.  One helper measured best (two or three
// were no faster, six slower); a report costs the solving work-group about a
// memory round trip, so reports are sparse and the lead is generous.
constexpr int NHELP = 1, HB = 4, HLEAD = 32, PROG_EVERY = 16;
constexpr int HSPIN = 4096, HSPIN_TOTAL = 1 << 16;

// Levels whose solved values are stored to z together.  A store joins the
// in-order count of outstanding memory operations that every later wait for
// a load drains, so one burst per ZB levels stalls at most once where a
// store per level could stall every level.  With the helper in place the
// measured effect is small (M^-T 603 -> 594 us, M^-1 unchanged at 705x321).
constexpr int ZB = 8;
static_assert(ZB % NBUF == 0 && PROG_EVERY % ZB == 0, "batch sizes");

// One triangular pass over the chunk's levels (ascending or descending),
// applying the diagonal inverse when DIAG.  NBUF operand sets rotate so that
// the loads of the next NBUF - 1 levels stay in flight while the current
// level computes.  The level count is rounded up to a multiple of ZB; the
// extra levels hold no valid cell.  Thread 0 reports the steps done to
// *prog (tag | steps) for the helpers, right after a level's loads, so that
// the report's store is not waited for before the next level.
template <int NV, int W, bool NINE, bool TR, bool ASC, bool DIAG>
__device__ __forceinline__ void cilu_pass(const float* R, int nlev, int njc, int ldi, size_t plane, int i0, int wk,
                                          int k, int c, const double* src, double* z, double (*ring)[NV][W],
                                          double (*tmp)[W], unsigned* prog, unsigned tag) {
  double f[NBUF][ct::nn<NINE>() * NV], d[NBUF][NV], x[NBUF];
  auto lev = [&](int step) { return ASC ? step : nlev - 1 - step; };
#pragma unroll
  for (int u = 0; u < NBUF - 1; ++u)
    cilu_fetch<NV, W, NINE, TR, ASC, DIAG>(R, nlev, njc, ldi, plane, i0, wk, k, c, lev(u), src, f[u], d[u], x[u]);
  for (int step = 0; step < nlev; step += ZB) {
    double zv[ZB];
#pragma unroll
    for (int u = 0; u < ZB; ++u) {
      const int cur = u % NBUF, nb = (u + NBUF - 1) % NBUF;
      cilu_fetch<NV, W, NINE, TR, ASC, DIAG>(R, nlev, njc, ldi, plane, i0, wk, k, c, lev(step + u + NBUF - 1), src,
                                             f[nb], d[nb], x[nb]);
      if (u == 0 && prog && threadIdx.x == 0 && (step % PROG_EVERY) == 0)
        __hip_atomic_store(prog, tag + (unsigned)step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      zv[u] = cilu_level<NV, W, NINE, ASC, DIAG>(lev(step + u), njc, wk, k, c, f[cur], d[cur], x[cur], ring, tmp);
    }
#pragma unroll
    for (int u = 0; u < ZB; ++u) {
      const int j = lev(step + u) - 2 * c;
      if (c < wk && j >= 0 && j < njc) z[k * plane + (size_t)j * ldi + i0 + c] = zv[u];
    }
  }
}

// Helper h of a chunk: loads one word of every 128-byte line of the slots a
// pass reads, level by level in the pass's order, steps h, h + NHELP, ...
// Pass p (0 ascending, 1 descending) of M^-1 reads the lower slots, then the
// upper slots and the diagonal; M^-T reads them the other way round.
template <int NV, int W, bool NINE, bool TR>
__device__ void cilu_helper(const float* R, int nlev, const unsigned* prog, unsigned epoch, int h, unsigned* sink) {
  constexpr int NT = NV * W, SLOTF = NV * NV * W, RS = cilu_rs<NV, W>(), NN = ct::nn<NINE>();
  constexpr int LPS = (SLOTF + 31) / 32;               // 128-byte lines per slot
  constexpr int LPT = ((NN + 1) * LPS + NT - 1) / NT;  // lines per thread and level
  float acc = 0.f;
  int budget = HSPIN_TOTAL;
  for (int pass = 0; pass < 2; ++pass) {
    const bool upper = (pass == 0) == TR;  // upper slots and the diagonal
    const int nslot = upper ? NN + 1 : NN;
    int off[LPT];
#pragma unroll
    for (int u = 0; u < LPT; ++u) {
      const int t = threadIdx.x + u * NT, si = t / LPS, ln = t % LPS;
      const int s = si >= nslot ? -1 : upper ? (si == NN ? 0 : ct::up<NINE>(si)) : ct::low<NINE>(si);
      off[u] = (s < 0 || ln * 32 >= SLOTF) ? -1 : s * SLOTF + ln * 32;
    }
    const unsigned base = (unsigned)(pass * nlev);
    for (int s0 = h; s0 < nlev; s0 += NHELP * HB) {
      const unsigned last = base + (unsigned)min(s0 + NHELP * (HB - 1), nlev - 1);
      for (int spin = 0; spin < HSPIN && budget > 0; ++spin, --budget) {
        const unsigned v = __hip_atomic_load(prog, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned done = (v >> 16) == epoch ? (v & 0xffffu) : 0u;
        if (done + HLEAD >= last) break;
        __builtin_amdgcn_s_sleep(4);
      }
      float v[HB][LPT];
#pragma unroll
      for (int b = 0; b < HB; ++b) {
        const int step = s0 + NHELP * b, L = pass == 0 ? step : nlev - 1 - step;
#pragma unroll
        for (int u = 0; u < LPT; ++u) v[b][u] = (step < nlev && off[u] >= 0) ? R[(size_t)L * RS + off[u]] : 0.f;
      }
#pragma unroll
      for (int b = 0; b < HB; ++b)
#pragma unroll
        for (int u = 0; u < LPT; ++u) acc += v[b][u];
    }
  }
  // keeps the loads: the sum is used, but it reaches memory only when a factor
  // is not finite, and then only a word that nothing reads
  if (acc != acc && threadIdx.x == 0) *sink = 1u;
}

// z = (LU)^-1 b, or (LU)^-T b when TR.  Thread (k, c): component k of the
// cell in column c.  Operands are loaded one level ahead of their use.
// Work-groups [0, nchunk) solve their chunk; with helpers (prog non-null),
// work-group npad (1 + m) + chunk is helper m of the chunk, npad being the
// chunk count rounded up to a multiple of 8, so it runs on the chunk's XCD.
template <int NV, int W, bool NINE, bool TR>
__global__ __launch_bounds__(NV * W) void cilu_apply_kernel(int nic, int njc, int ldi, int nlev, const float* rec,
                                                            size_t tcopy, const double* b, double* z, unsigned* prog,
                                                            unsigned epoch) {
  constexpr int RS = cilu_rs<NV, W>();
  // column index fastest: the lanes of a wave are consecutive columns, so
  // their LDS reads and writes fall in distinct banks
  __shared__ double ring[8][NV][W];
  __shared__ double tmp[NV][W];
  const int nchunk = (nic + W - 1) / W, npad = (nchunk + 7) & ~7;
  const int chunk = (int)blockIdx.x % npad, role = (int)blockIdx.x / npad;
  if (chunk >= nchunk) return;
  const float* R = rec + (TR ? tcopy : 0) + (size_t)chunk * nlev * RS;
  if (role > 0) {
    cilu_helper<NV, W, NINE, TR>(R, nlev, prog + chunk, epoch, role - 1, prog + npad);
    return;
  }
  const int i0 = chunk * W, wk = min(W, nic - i0);
  const size_t plane = (size_t)njc * ldi;
  const int k = threadIdx.x / W, c = threadIdx.x % W;
  unsigned* pc = prog ? prog + chunk : nullptr;
  const unsigned tag = epoch << 16;
  // ring cells never written hold zeros (their f is zero too)
  for (int e = threadIdx.x; e < W * 8 * NV; e += NV * W) (&ring[0][0][0])[e] = 0.0;
  __syncthreads();
  // non-transposed: unit-lower L ascending, then D + U descending;
  // transposed: (D + U)^T ascending, then unit-lower L^T descending
  cilu_pass<NV, W, NINE, TR, true, TR>(R, nlev, njc, ldi, plane, i0, wk, k, c, b, z, ring, tmp, pc, tag);
  // the second pass reads what the first stored: complete every store first
  __syncthreads();
  cilu_pass<NV, W, NINE, TR, false, !TR>(R, nlev, njc, ldi, plane, i0, wk, k, c, z, z, ring, tmp, pc,
                                         tag + (unsigned)nlev);
}

// progress words of the helpers: one per chunk (padded to a multiple of 8)
// and one sink, after both copies
static size_t cilu_prog_words(int nic, int w) { return (size_t)(((nic + w - 1) / w + 7) & ~7) + 1; }

// doubles before the gathered factorisation operands: both copies of FP32
// factors, then the progress words
static size_t cilu_ops_offset(int nic, int njc, int nv, int w) {
  return cilu_copy_floats(nic, njc, nv, w) + (cilu_prog_words(nic, w) + 1) / 2;
}

// in doubles of storage: the factors, the progress words and the gathered
// operands (as many doubles as one copy has floats)
size_t cilu_size(int nic, int njc, int nv, int w) {
  return cilu_ops_offset(nic, njc, nv, w) + cilu_copy_floats(nic, njc, nv, w);
}

template <int NV, int W, bool NINE>
static void cilu_factor_t(int nic, int njc, int ldi, const double* idt, const double* blocks, double* store,
                          hipStream_t s) {
  const int nchunk = (nic + W - 1) / W, nlev = njc + 2 * W - 2;
  const size_t total = cilu_copy_floats(nic, njc, NV, W);
  double* ops = store + cilu_ops_offset(nic, njc, NV, W);
  const size_t nblk = std::min<size_t>((total + 255) / 256, (size_t)1 << 20);
  hipLaunchKernelGGL((cilu_gather_kernel<NV, W>), dim3((unsigned)nblk), dim3(256), 0, s, nic, njc, ldi, nlev, idt,
                     blocks, ops, total);
  hipLaunchKernelGGL((cilu_factor_kernel<NV, W, NINE>), dim3(nchunk), dim3(W * NV * NV), 0, s, nic, njc, nlev, ops,
                     reinterpret_cast<float*>(store), total);
}

// Launch tag of an apply: helpers only trust progress words that carry it.
static std::atomic<unsigned> g_apply_epoch{0};

template <int NV, int W, bool NINE>
static void cilu_apply_t(int nic, int njc, int ldi, double* store, const double* b, double* z, bool tr,
                         hipStream_t s) {
  const int nchunk = (nic + W - 1) / W, nlev = njc + 2 * W - 2, npad = (nchunk + 7) & ~7;
  const size_t tcopy = cilu_copy_floats(nic, njc, NV, W);
  const float* rec = reinterpret_cast<const float*>(store);
  // the progress words follow the two factor copies; the apply writes them
  unsigned* prog = reinterpret_cast<unsigned*>(reinterpret_cast<float*>(store) + 2 * tcopy);
  // the progress word holds tag << 16 | steps; without room for 2 nlev steps, no helpers
  const bool helpers = 2 * nlev < (1 << 16);
  unsigned epoch = (g_apply_epoch.fetch_add(1) % 0xffffu) + 1;  // never 0, the value of fresh memory
  const dim3 grid(helpers ? npad * (1 + NHELP) : nchunk);
  if (!helpers) prog = nullptr, epoch = 0;
  if (tr)
    hipLaunchKernelGGL((cilu_apply_kernel<NV, W, NINE, true>), grid, dim3(NV * W), 0, s, nic, njc, ldi, nlev, rec,
                       tcopy, b, z, prog, epoch);
  else
    hipLaunchKernelGGL((cilu_apply_kernel<NV, W, NINE, false>), grid, dim3(NV * W), 0, s, nic, njc, ldi, nlev, rec,
                       tcopy, b, z, prog, epoch);
}

bool cilu_width_ok(int w) { return w == 8 || w == 16 || w == 24 || w == 32; }

template <bool NINE>
static void cilu_factor_n(int nic, int njc, int ldi, int nv, int w, const double* idt, const double* blocks,
                          double* store, hipStream_t s) {
  if (nv == 5) {
    if (w == 8) cilu_factor_t<5, 8, NINE>(nic, njc, ldi, idt, blocks, store, s);
    else if (w == 16) cilu_factor_t<5, 16, NINE>(nic, njc, ldi, idt, blocks, store, s);
    else if (w == 24) cilu_factor_t<5, 24, NINE>(nic, njc, ldi, idt, blocks, store, s);
    else cilu_factor_t<5, 32, NINE>(nic, njc, ldi, idt, blocks, store, s);
  } else {
    if (w == 8) cilu_factor_t<4, 8, NINE>(nic, njc, ldi, idt, blocks, store, s);
    else if (w == 16) cilu_factor_t<4, 16, NINE>(nic, njc, ldi, idt, blocks, store, s);
    else if (w == 24) cilu_factor_t<4, 24, NINE>(nic, njc, ldi, idt, blocks, store, s);
    else cilu_factor_t<4, 32, NINE>(nic, njc, ldi, idt, blocks, store, s);
  }
}

template <bool NINE>
static void cilu_apply_n(int nic, int njc, int ldi, int nv, int w, double* store, const double* b, double* z,
                         bool transpose, hipStream_t s) {
  if (nv == 5) {
    if (w == 8) cilu_apply_t<5, 8, NINE>(nic, njc, ldi, store, b, z, transpose, s);
    else if (w == 16) cilu_apply_t<5, 16, NINE>(nic, njc, ldi, store, b, z, transpose, s);
    else if (w == 24) cilu_apply_t<5, 24, NINE>(nic, njc, ldi, store, b, z, transpose, s);
    else cilu_apply_t<5, 32, NINE>(nic, njc, ldi, store, b, z, transpose, s);
  } else {
    if (w == 8) cilu_apply_t<4, 8, NINE>(nic, njc, ldi, store, b, z, transpose, s);
    else if (w == 16) cilu_apply_t<4, 16, NINE>(nic, njc, ldi, store, b, z, transpose, s);
    else if (w == 24) cilu_apply_t<4, 24, NINE>(nic, njc, ldi, store, b, z, transpose, s);
    else cilu_apply_t<4, 32, NINE>(nic, njc, ldi, store, b, z, transpose, s);
  }
}

hipError_t cilu_factor(int nic, int njc, int ldi, int nv, int w, bool nine, const double* idt, const double* blocks,
                       double* store, hipStream_t s) {
  if (nine) cilu_factor_n<true>(nic, njc, ldi, nv, w, idt, blocks, store, s);
  else cilu_factor_n<false>(nic, njc, ldi, nv, w, idt, blocks, store, s);
  return hipGetLastError();
}

hipError_t cilu_apply(int nic, int njc, int ldi, int nv, int w, bool nine, double* store, const double* b, double* z,
                      bool transpose, hipStream_t s) {
  if (nine) cilu_apply_n<true>(nic, njc, ldi, nv, w, store, b, z, transpose, s);
  else cilu_apply_n<false>(nic, njc, ldi, nv, w, store, b, z, transpose, s);
  return hipGetLastError();
}

}  // namespace sg
