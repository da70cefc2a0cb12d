// d4_gpulu.hip -- the host simulation's linear solves as a dense LU on the GPU,
// in place of UMFPACK's umfpack_di_symbolic / _numeric / _solve
// (simulation.cpp:635-659), for one matrix (GpuDenseLU) or for the matrices of
// many realizations at once (BatchLU).
//
// At the sizes of a realization (N <= a few thousand) one dense LU on the GPU
// takes milliseconds where a sparse LU of the cutoff pattern on a host core
// takes up to seconds (the 2D cutoff graph fills in heavily), so this is the
// solver the batched ensemble and large single runs use.  A matrix given in
// CSC form is factored in a symmetric order taken from its pattern (reverse
// Cuthill-McKee when that halves the work, see band_order in d4_solver.h),
// and only the blocks of its envelope are filled, factored and read.
//
// Layout: matrix k is column-major of order npad_k = gpu_lu_padded(n_k) (n_k
// rounded up to 64, the factorization's block), the CSC entries scattered into its leading n_k x n_k
// block and the identity beyond it.  The padding changes nothing in the
// solution: its rows and columns are decoupled from the leading block and
// never enter a pivot choice.  It gives every matrix an order that is a
// multiple of the factorization's block.  A matrix's factors do not depend
// on the batch it is in (bitwise; the ensemble tests check it).
// GpuDenseLU is BatchLU with K = 1.
//
// Pivoting: threshold partial pivoting with a preference for the diagonal and
// u = 0.001, the diagonal tolerance of UMFPACK's symmetric strategy
// (Control[UMFPACK_SYM_PIVOT_TOLERANCE]); UMFPACK, the reference's solver,
// selects that strategy by default for a symmetric nonzero pattern with a
// nonzero diagonal, which these step matrices have.  Every matrix is first
// factored without row interchanges by this file's blocked LU
// (lu_diag_kernel, lu_panel_kernel, lu_update_kernel: three launches per 64
// columns for the whole ragged batch, the trailing update on the FP64 matrix
// cores); if every pivot is finite and nonzero and every multiplier
// |l_rc| <= 1/u = 1000, the diagonal was an acceptable pivot at every stage
// and the factorization stands.  Otherwise the matrix is scattered again
// and factored with partial pivoting (rocSOLVER getrf, one call per matrix).  The choice
// depends on the matrix values only.
//
// Solves: a triangular-solve kernel of this file (lu_solve_kernel), one
// workgroup per matrix for all matrices of the batch in one launch, in place
// of getrs (whose single-right-hand-side solves are launch bound: 1-5 ms each
// at N = 256..4096).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <set>
#include <mutex>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/d4_sim.h"
#include "d4_ctx.h"
#include "d4_solver.h"

namespace d4host {

int gpu_lu_padded(int n) { return n <= 0 ? 0 : (n + 63) / 64 * 64; }

namespace {

// one stream + rocBLAS handle per (thread, device); kept until the process ends
struct ThreadRes {
    hipStream_t stream = nullptr;
    rocblas_handle handle = nullptr;
};

ThreadRes* thread_res(int device)
{
    if (t_host_worker) return nullptr;  // no HIP from the ensemble's host workers (d4_solver.h)
    static std::mutex m;
    static std::map<std::pair<std::thread::id, int>, ThreadRes*> reg;
    std::lock_guard<std::mutex> l(m);
    auto key = std::make_pair(std::this_thread::get_id(), device);
    auto it = reg.find(key);
    if (it != reg.end()) return it->second;
    if (hipSetDevice(device) != hipSuccess) return nullptr;
    ThreadRes* r = new ThreadRes();
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess ||
        rocblas_create_handle(&r->handle) != rocblas_status_success ||
        rocblas_set_stream(r->handle, r->stream) != rocblas_status_success) {
        delete r;
        return nullptr;
    }
    reg[key] = r;
    return r;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    bool ensure(size_t count)
    {
        if (count <= cap) return true;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = count;
        return true;
    }
};

template <class T>
struct PinBuf {
    T* p = nullptr;
    size_t cap = 0;
    ~PinBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    bool ensure(size_t count)
    {
        if (count <= cap) return true;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        if (hipHostMalloc(&p, sizeof(T) * std::max<size_t>(count, 1), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return false;
        }
        cap = count;
        return true;
    }
};

// Symmetric order and block envelope of a matrix (aux, per matrix): at
// ord_off, ord[new] = old then inv[old] = new (n entries each); at env_off,
// per block b of nb = npad / kFB: last[b], then top[b], then first[b] (see
// block_envelope).  -1: the natural order / the whole matrix.
struct MatDesc {
    long long a_off;  // first entry of the matrix in the dense store
    int n, npad;
    int ipiv_off;  // its npad pivots (matrices refactored with pivoting)
    int ord_off, env_off;
    int row_base;  // rows stored as r + row_base in ai
    const int* ap;  // its CSC on the device: column c at ai/ax[ap[c] .. ap[c + 1])
    const int* ai;
    const double* ax;
};

struct SolveDesc {
    long long a_off;
    int n, npad;
    int b_off;     // its right-hand side in the packed vector (n entries)
    int perm_off;  // its row permutation, or -1 when rows stayed in place
    int ord_off, env_off;
};

constexpr double kMaxMultiplier = 1000.0;  // threshold partial pivoting with u = 0.001
constexpr int kFB = 64;                    // block order of the factorization and of the envelope
constexpr int kSB = kFB;                   // block of the triangular solves
constexpr int kSolveThreads = 1024;
constexpr int kSolveLds = 2 * kSB * kSB + 2 * kSB;  // doubles of LDS before x: two staged blocks, two diagonals

// Column groups of the per-column kernels below: a 64-column block of a matrix is split over `split`
// workgroups (a power of two up to 16) when a launch would otherwise hold fewer than ~2048 of them
inline int column_split(long long blocks)
{
    int s = 1;
    while (s < 16 && blocks * s < 2048) s *= 2;
    return s;
}

// identity beyond n, zero in the leading block: the columns of group blockIdx.x % split of block
// blockIdx.x / split of matrix blockIdx.y, one wavefront per column at a time.  With an envelope, only
// the rows the factorization, the check and the solves touch in these columns: blocks top .. last of
// their block.
__global__ __launch_bounds__(256) void fill_identity_kernel(const MatDesc* __restrict__ desc, double* __restrict__ A,
                                                            const int* __restrict__ aux, int split)
{
    const MatDesc m = desc[blockIdx.y];
    const int b = blockIdx.x / split, cw = kFB / split, c0 = b * kFB + (blockIdx.x % split) * cw;
    if (b * kFB >= m.npad) return;
    int r0 = 0, r1 = m.npad;
    if (m.env_off >= 0) {
        const int nb = m.npad / kFB;
        r0 = aux[m.env_off + nb + b] * kFB;
        r1 = min(m.npad, (aux[m.env_off + b] + 1) * kFB);
    }
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int cc = w; cc < cw; cc += 4) {
        const int c = c0 + cc;
        double* col = A + m.a_off + (size_t)c * m.npad;
        const double diag = c >= m.n ? 1.0 : 0.0;
        for (int r = r0 + lane; r < r1; r += 64) col[r] = r == c ? diag : 0.0;
    }
}

// the CSC entries of column c (one thread) into the dense matrix, rows and
// columns renumbered by inv when the matrix is reordered
__global__ void scatter_kernel(const MatDesc* __restrict__ desc, double* __restrict__ A, const int* __restrict__ aux)
{
    const MatDesc m = desc[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m.n) return;
    const int* inv = m.ord_off >= 0 ? aux + m.ord_off + m.n : nullptr;
    double* col = A + m.a_off + (size_t)(inv ? inv[c] : c) * m.npad;
    for (int q = m.ap[c]; q < m.ap[c + 1]; ++q) {
        const int r = m.ai[q] - m.row_base;
        col[inv ? inv[r] : r] = m.ax[q];
    }
}

// one dense matrix already on the device into the padded store: element (r, c)
// of the source at src[r * rs + c * cs]; identity beyond n
__global__ void dense_copy_kernel(const MatDesc* __restrict__ desc, double* __restrict__ A,
                                  const double* __restrict__ src, long long rs, long long cs)
{
    const MatDesc m = desc[blockIdx.y];
    const int c = blockIdx.x;
    if (c >= m.npad) return;
    double* col = A + m.a_off + (size_t)c * m.npad;
    for (int r = threadIdx.x; r < m.npad; r += blockDim.x)
        col[r] = (r < m.n && c < m.n) ? src[r * rs + c * cs] : (r == c ? 1.0 : 0.0);
}

// after the no-pivot factorization: flag[k] = 1 unless every pivot is finite
// and nonzero and every multiplier |l_rc| <= kMaxMultiplier, i.e. unless
// threshold partial pivoting with u = 1/kMaxMultiplier would keep the diagonal
// at every stage
// (with an envelope the multipliers below its last block are zero, unread).
// The columns of group blockIdx.x % split of block blockIdx.x / split of matrix blockIdx.y, one wavefront
// per column at a time.
__global__ __launch_bounds__(256) void multiplier_check_kernel(const MatDesc* __restrict__ desc,
                                                               const double* __restrict__ A, int* __restrict__ flag,
                                                               const int* __restrict__ aux, int split)
{
    const MatDesc m = desc[blockIdx.y];
    const int b = blockIdx.x / split, cw = kFB / split, c0 = b * kFB + (blockIdx.x % split) * cw;
    if (c0 >= m.n) return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r1 = m.env_off >= 0 ? min(m.n, (aux[m.env_off + b] + 1) * kFB) : m.n;
    bool bad = false;
    for (int cc = w; cc < cw; cc += 4) {
        const int c = c0 + cc;
        if (c >= m.n) break;
        const double* col = A + m.a_off + (size_t)c * m.npad;
        if (lane == 0) {
            const double p = std::fabs(col[c]);
            bad |= !(p > 0.0) || !(p <= 1.7976931348623157e308);
        }
        for (int r = c + 1 + lane; r < r1; r += 64) bad |= !(std::fabs(col[r]) <= kMaxMultiplier);
    }
    if (__any(bad) && lane == 0) atomicOr(flag + blockIdx.y, 1);
}

// LAPACK pivots (row i swapped with row ipiv[i] - 1, in order) into the
// permutation perm (solution row i comes from right-hand-side row perm[i]),
// in place; one workgroup per matrix, the swaps by one thread in LDS (kLds) or
// in global memory
template <bool kLds>
__global__ void pivots_to_perm_kernel(const MatDesc* __restrict__ desc, int* __restrict__ ipiv,
                                      int* __restrict__ scratch)
{
    extern __shared__ int sp[];
    const MatDesc m = desc[blockIdx.x];
    int* iv = ipiv + m.ipiv_off;
    int* p = kLds ? sp : scratch + 2 * (size_t)m.ipiv_off;
    int* q = p + m.npad;
    for (int i = threadIdx.x; i < m.npad; i += blockDim.x) {
        p[i] = i;
        const int j = iv[i] - 1;
        q[i] = (j >= i && j < m.npad) ? j : i;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < m.npad; ++i) {
            const int j = q[i];
            if (j != i) {
                const int t = p[i];
                p[i] = p[j];
                p[j] = t;
            }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < m.npad; i += blockDim.x) iv[i] = p[i];
}

__device__ inline double lane_value(double v, int lane)  // v of `lane` (uniform) in every lane
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// x := U^{-1} L^{-1} P b for the factors in A (getrf's layout: unit lower L
// below the diagonal, U on and above), one workgroup per matrix.  Blocked by
// kSB rows: the diagonal block is staged in LDS and solved by wave 0 (lane l
// owns row l, the solved value of lane c broadcast at step c), then the
// other rows are updated with the block's columns,
// x_r -= sum_c a_rc x_c.  For npad <= 1024 a row is shared by four threads:
// thread k of the row accumulates the columns c = k, k + 4, ... in ascending
// order, and the four partial sums are added as (s0 + s1) + (s2 + s3) in
// every thread of the row; larger matrices sum a row in one thread, c
// ascending.  The order of operations
// is fixed by the matrix alone.  x lives in LDS (kLds) or in global scratch;
// the arithmetic is the same either way.  A reordered matrix gathers b
// through ord and scatters x back through it; with an envelope the updates
// stop at its bounds, past which the factors' columns are zero (a zero sum
// leaves x_r as it is, so the bits are those of the full sweep); a matrix
// whose full flag is set is swept whole.
// Latency: with four threads a row, every thread loads its first row's 16
// factor entries of the block before the block's diagonal solve, which they
// do not depend on, so that the loads are in flight during it (one thread a
// row loads the row's 64 entries 32 at a time after it); the next diagonal
// block is staged (by waves 1..15, into the other of two LDS buffers) at the
// same time; wave 0 reads its entries of the block kAhead steps before their
// use.
template <bool kLds>
__global__ __launch_bounds__(kSolveThreads) void lu_solve_kernel(const SolveDesc* __restrict__ desc,
                                                                  const double* __restrict__ A,
                                                                  const int* __restrict__ perm, double* __restrict__ b,
                                                                  double* __restrict__ scratch,
                                                                  const int* __restrict__ aux,
                                                                  const int* __restrict__ full)
{
    constexpr int kAhead = 8, kPre = 16;
    constexpr int kStage = (kSB * kSB + kSolveThreads - 65) / (kSolveThreads - 64);  // entries per staging thread
    extern __shared__ double lds[];
    const SolveDesc m = desc[blockIdx.x];
    const int n = m.n, tid = threadIdx.x;
    const size_t lda = (size_t)m.npad;
    const double* Am = A + m.a_off;
    double* bm = b + m.b_off;
    double* Dbuf = lds;                   // two kSB x kSB diagonal blocks, column-major
    double* Ddiag = lds + 2 * kSB * kSB;  // the diagonals of U's blocks staged there
    double* x = kLds ? lds + kSolveLds : scratch + m.b_off;
    const int* ord = m.ord_off >= 0 ? aux + m.ord_off : nullptr;
    // full[k] set: matrix k was refactored with pivoting, whose factors are not held to its envelope
    const int* env = (m.env_off >= 0 && !(full && full[blockIdx.x])) ? aux + m.env_off : nullptr;
    const int nb = m.npad / kSB;
    const int nblk = (n + kSB - 1) / kSB;
    // diagonal block number s of the two sweeps (0..2 nblk - 1) into buffer s & 1, by the threads t0,
    // t0 + stride, ...: the strictly lower part of L's block (first sweep) or the strictly upper part of U's
    // block and its diagonal (second sweep), zeros elsewhere
    auto stage = [&](int s, int t0, int stride) {
        const bool up = s >= nblk;
        const int bi = up ? 2 * nblk - 1 - s : s;
        const int jb = bi * kSB, B = min(kSB, n - jb);
        double* D = Dbuf + (s & 1) * kSB * kSB;
        // all of a thread's loads are issued before its first store (at most kStage entries a thread)
        double v[kStage];
#pragma unroll
        for (int i = 0; i < kStage; ++i) {
            const int e = t0 + i * stride, r = e % kSB, c = e / kSB;
            v[i] = (e < kSB * kSB && r < B && c < B) ? Am[(size_t)(jb + c) * lda + jb + r] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kStage; ++i) {
            const int e = t0 + i * stride, r = e % kSB, c = e / kSB;
            if (e < kSB * kSB) {
                D[e] = (up ? r < c : r > c) ? v[i] : 0.0;
                if (up && r == c) Ddiag[(s & 1) * kSB + r] = (r < B) ? v[i] : 1.0;
            }
        }
    };
    for (int i = tid; i < n; i += kSolveThreads) {
        int p = m.perm_off >= 0 ? perm[m.perm_off + i] : i;  // padded rows never move into [0, n)
        p = (p >= 0 && p < n) ? p : i;
        x[i] = bm[ord ? ord[p] : p];
    }
    if (nblk > 0) stage(0, tid, kSolveThreads);
    __syncthreads();

    // npad <= 1024: each row update shared by four threads of waves 1..15 (wave 0 solves the diagonal
    // blocks), thread k of a row taking the columns k, k + 4, ..., its 16 entries of the row loaded before
    // the diagonal solve; larger: one thread of the whole workgroup per row, its 64 entries 32 at a time
    const bool split = m.npad <= 1024;
    const int ut = tid - 64;
    const int k = split && ut >= 0 ? ut & 3 : 0;
    const int q = split ? (ut >= 0 ? ut >> 2 : (1 << 29)) : tid;
    const int rows = split ? (kSolveThreads - 64) / 4 : kSolveThreads;
    for (int s = 0; s < 2 * nblk; ++s) {
        const int pass = s < nblk ? 0 : 1;  // 0: L y = P b top down, 1: U x = y bottom up
        const int bi = pass == 0 ? s : 2 * nblk - 1 - s;
        const int jb = bi * kSB, B = min(kSB, n - jb);
        const double* D = Dbuf + (s & 1) * kSB * kSB;
        const int r0 = pass == 0 ? jb + B : (env ? env[2 * nb + bi] * kSB : 0);
        const int r1 = pass == 0 ? (env ? min(n, (env[bi] + 1) * kSB) : n) : jb;
        const bool whole = B == kSB;
        const int rf = r0 + q;
        double av[kPre];  // split: the first row's 16 entries
        if (split && whole && rf < r1) {
            const double* a = Am + (size_t)(jb + k) * lda + rf;
#pragma unroll
            for (int c = 0; c < kPre; ++c) av[c] = a[(size_t)(4 * c) * lda];
        }
        if (tid < 64) {
            // lane l owns row l; the staged block is zero outside the triangle a step updates, so every
            // lane takes every step with no mask (a zero factor leaves v as it is)
            const int l = tid;
            double v = l < B ? x[jb + l] : 0.0;
            double d[kSB];  // d[c] = D(l, c)
            if (pass == 0) {
#pragma unroll
                for (int kk = 0; kk < kAhead; ++kk) d[kk] = D[kk * kSB + l];
#pragma unroll
                for (int c = 0; c < kSB - 1; ++c) {
                    if (c + kAhead < kSB - 1) d[c + kAhead] = D[(c + kAhead) * kSB + l];
                    v = fma(-d[c], lane_value(v, c), v);
                }
            } else {
                // x_c = v_c / u_cc, correctly rounded, from the reciprocal: q = v r, then q + (v - q u) r
                // (exact residual by fma) is the quotient IEEE division gives (away from the overflow and
                // underflow thresholds), with no division on the step-to-step chain
                const double ul = Ddiag[(s & 1) * kSB + l], rl = 1.0 / ul;
#pragma unroll
                for (int kk = 0; kk < kAhead; ++kk) d[kSB - 1 - kk] = D[(kSB - 1 - kk) * kSB + l];
#pragma unroll
                for (int c = kSB - 1; c >= 0; --c) {
                    if (c - kAhead >= 0) d[c - kAhead] = D[(c - kAhead) * kSB + l];
                    if (c < B) {
                        const double vc = lane_value(v, c), uc = lane_value(ul, c), rc = lane_value(rl, c);
                        const double q0 = vc * rc;
                        v = fma(-d[c], fma(fma(-q0, uc, vc), rc, q0), v);
                    }
                }
                v = v / ul;  // lane l's own quotient, as step l formed it (no later step changes v_l)
            }
            if (l < B) x[jb + l] = v;
        } else if (s + 1 < 2 * nblk) {
            stage(s + 1, tid - 64, kSolveThreads - 64);
        }
        __syncthreads();
        if (split) {
            // rows r0 + q, r0 + q + rows, ...: the same rows for the four threads of a row
            for (int r = rf; r < r1; r += rows) {
                const double* a = Am + (size_t)(jb + k) * lda + r;
                double acc = 0.0;
                if (whole) {
                    if (r != rf) {
#pragma unroll
                        for (int c = 0; c < kPre; ++c) av[c] = a[(size_t)(4 * c) * lda];
                    }
#pragma unroll
                    for (int c = 0; c < kPre; ++c) acc = fma(av[c], x[jb + k + 4 * c], acc);
                } else {
                    for (int c = k; c < B; c += 4) acc = fma(a[(size_t)(c - k) * lda], x[jb + c], acc);
                }
                acc += __shfl_xor(acc, 1);
                acc += __shfl_xor(acc, 2);
                if (k == 0) x[r] -= acc;
            }
        } else {
            for (int r = rf; r < r1; r += rows) {
                const double* a = Am + (size_t)jb * lda + r;
                double acc = 0.0;
                if (whole) {
#pragma unroll
                    for (int h = 0; h < kSB; h += 2 * kPre) {
                        double bv[2 * kPre];
#pragma unroll
                        for (int c = 0; c < 2 * kPre; ++c) bv[c] = a[(size_t)(h + c) * lda];
#pragma unroll
                        for (int c = 0; c < 2 * kPre; ++c) acc = fma(bv[c], x[jb + h + c], acc);
                    }
                } else {
                    for (int c = 0; c < B; ++c) acc = fma(a[(size_t)c * lda], x[jb + c], acc);
                }
                x[r] -= acc;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += kSolveThreads) bm[ord ? ord[i] : i] = x[i];
}

// ---------------------------------------------------------------------------
// Blocked right-looking LU without row interchanges for a ragged batch (every
// matrix of order npad, a multiple of kFB).  Step p (j0 = kFB p) covers the
// matrices with npad > j0, a prefix of the batch sorted by npad, largest first:
//   lu_diag_kernel    factors the diagonal block A11 = L11 U11 (one workgroup per matrix);
//   lu_panel_kernel   L21 = A21 U11^-1 and U12 = L11^-1 A12, one wavefront per kFB x kFB block;
//   lu_update_kernel  A22 -= L21 U12, one workgroup per kFB x kFB tile, on the matrix cores.
// Per step and matrix, tb[i] = T: the blocks p + 1 .. p + T below and right
// of the diagonal block are the ones that can be nonzero (all trailing blocks,
// or those within the matrix's envelope); the panels and the update cover
// those.  A block of L21 or U12 past the envelope is zero, and an update by
// zero leaves an entry as it is, so skipping them changes no bits.
// The order of operations for an entry depends on its matrix alone (the
// workgroups of other matrices only add parallelism), so a matrix gets the
// same factors in any batch.
constexpr int kFLd = kFB + 1;  // LDS column stride (odd: conflict-free column and row walks)

struct FacDesc {
    long long a_off;
    int npad;
    int pad_;
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Work item of this workgroup such that each of the 8 XCDs (which receive the
// workgroups of a launch round-robin) runs a contiguous range of items: the
// tiles of one block row or column then share that XCD's L2.  A bijection on
// [0, gridDim.x) whatever the dispatch order.
__device__ inline int xcd_item()
{
    const int n = gridDim.x, x = blockIdx.x, xcd = x & 7, idx = x >> 3;
    const int q = n >> 3, r = n & 7;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// index of the matrix owning work item g: pre[i] <= g < pre[i + 1], pre[0] = 0
__device__ inline int owner(const int* __restrict__ pre, int count, int g)
{
    int lo = 0, hi = count;  // pre[lo] <= g < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// LU without interchanges of each matrix's diagonal block, one wavefront per
// matrix: lane r holds row r in registers.  At step j every lane reads the
// pivot row (lane j) with readlane, and each row below forms its multiplier
// l = a_rj / a_jj and updates a_rc -= l a_jc (c > j); rows at or above the
// pivot take l = 0, which leaves them as they are.  No barrier and no LDS.
__global__ __launch_bounds__(64) void lu_diag_kernel(const FacDesc* __restrict__ fd, double* __restrict__ A, int j0)
{
    const FacDesc m = fd[blockIdx.x];
    const size_t lda = (size_t)m.npad;
    double* blk = A + m.a_off + (size_t)j0 * lda + j0;
    const int r = threadIdx.x;
    double a[kFB];
#pragma unroll
    for (int c = 0; c < kFB; ++c) a[c] = blk[(size_t)c * lda + r];
#pragma unroll
    for (int j = 0; j < kFB - 1; ++j) {
        const double l = r > j ? a[j] / lane_value(a[j], j) : 0.0;
        if (r > j) a[j] = l;
#pragma unroll
        for (int c = j + 1; c < kFB; ++c) a[c] = fma(-l, lane_value(a[c], j), a[c]);
    }
#pragma unroll
    for (int c = 0; c < kFB; ++c) blk[(size_t)c * lda + r] = a[c];
}

// f(std::integral_constant<int, 0>) ... f(std::integral_constant<int, N - 1>), in order: a loop whose
// index is a constant expression in its body (a DPP lane select must be one)
template <int... I, class Fn>
__device__ inline void static_for_seq(std::integer_sequence<int, I...>, Fn&& f)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class Fn>
__device__ inline void static_for(Fn&& f)
{
    static_for_seq(std::make_integer_sequence<int, N>{}, f);
}

// v of lane 16 g + L, in every lane of the 16-lane row g (DPP row_newbcast)
template <int L>
__device__ inline double row_value(double v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + L, 0xf, 0xf, true);
}

// acc = fma(-f', x, acc) with f' the f of lane 16 g + L in the lanes of row g: one v_fmac_f64 whose first
// operand crosses the row by DPP row_newbcast (no readlane, no SGPR round trip).  A DPP source must not be
// written by a VALU instruction in the two before it, a hazard the compiler does not see through inline
// asm: f is always a value loaded from LDS several instructions earlier (the build's assembly is checked
// for it by tests/test_lu_isa_cpu.py).
template <int L>
__device__ inline void fmac_row(double& acc, double f, double x)
{
    asm("v_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf" : "+v"(acc) : "v"(f), "v"(x), "i"(L));
}

// One 64 x 64 block per wavefront, four blocks of one matrix per workgroup
// (each matrix's items padded to a multiple of four), the diagonal block
// staged in LDS once; after that barrier no wavefront waits on another.  In
// an L21 block lane lo holds row lo of X in registers (X U11 = B, columns left
// to right), in a U12 block column lo (L11 X = B, rows top to bottom).  At
// step c the factor's row c of U11 (or column c of L11) is read from LDS into
// four registers, lane 16 g + L holding entry 16 k + L of register k in each
// of the four rows g, one step ahead; each update is then one v_fmac_f64 that
// takes its factor entry across the row by DPP.  The order of operations of
// every entry is that of the substitution written out: x_c2 -= x_c u_{c, c2},
// c ascending (fma(-u, x, y) and fma(-x, u, y) round alike).
__global__ __launch_bounds__(256, 1) void lu_panel_kernel(const FacDesc* __restrict__ fd, const int* __restrict__ pre,
                                                          const int* __restrict__ tb, int count,
                                                          double* __restrict__ A, int j0)
{
    __shared__ double Dg[kFB * kFLd];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g0 = xcd_item() * 4;  // the workgroup's first item; its four items share a matrix
    const int i = owner(pre, count, g0);
    const FacDesc m = fd[i];
    const size_t lda = (size_t)m.npad;
    const int T = tb[i];  // blocks below and right of the diagonal block that can be nonzero
    const double* dg = A + m.a_off + (size_t)j0 * lda + j0;
    const int t = threadIdx.x, lo = t & 63, q = t >> 6;
    for (int c = q; c < kFB; c += 4) Dg[c * kFLd + lo] = dg[(size_t)c * lda + lo];
    __syncthreads();
    const int item = g0 - pre[i] + wv;
    if (item >= 2 * T) return;                     // padding
    const bool rowblk = item < T;                  // L21 block (below) or U12 block (right)
    const int b = (rowblk ? item : item - T) + 1;  // block index past the diagonal
    const int L = lo & 15;
    double x[kFB], F[2][4];  // F[s & 1][k]: step s's factor entries 16 k + L
    if (rowblk) {
        double* blk = A + m.a_off + (size_t)j0 * lda + (j0 + kFB * b);
        double R[4];  // lane 16 g + L of R[k]: 1 / U11(16 k + L, 16 k + L)
#pragma unroll
        for (int k = 0; k < 4; ++k) R[k] = 1.0 / Dg[(16 * k + L) * kFLd + 16 * k + L];
#pragma unroll
        for (int c = 0; c < kFB; ++c) x[c] = blk[(size_t)c * lda + lo];
#pragma unroll
        for (int k = 0; k < 4; ++k) F[0][k] = Dg[(16 * k + L) * kFLd];  // U11(0, 16 k + L)
        static_for<kFB>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            if constexpr (c + 1 < kFB) {
#pragma unroll
                for (int k = (c + 2) >> 4; k < 4; ++k) F[(c + 1) & 1][k] = Dg[(16 * k + L) * kFLd + c + 1];
            }
            __builtin_amdgcn_sched_barrier(0);  // the next step's loads stay ahead of this step's updates
            x[c] *= row_value<c & 15>(R[c >> 4]);
            static_for<kFB>([&](auto cc2) {
                constexpr int c2 = decltype(cc2)::value;
                if constexpr (c2 > c) fmac_row<c2 & 15>(x[c2], F[c & 1][c2 >> 4], x[c]);  // -= U11(c, c2) x_c
            });
        });
#pragma unroll
        for (int c = 0; c < kFB; ++c) blk[(size_t)c * lda + lo] = x[c];
    } else {
        double* col = A + m.a_off + (size_t)(j0 + kFB * b + lo) * lda + j0;
#pragma unroll
        for (int r = 0; r < kFB; ++r) x[r] = col[r];
#pragma unroll
        for (int k = 0; k < 4; ++k) F[0][k] = Dg[16 * k + L];  // L11(16 k + L, 0)
        static_for<kFB - 1>([&](auto rr) {
            constexpr int r = decltype(rr)::value;
            if constexpr (r + 1 < kFB - 1) {
#pragma unroll
                for (int k = (r + 2) >> 4; k < 4; ++k) F[(r + 1) & 1][k] = Dg[(r + 1) * kFLd + 16 * k + L];
            }
            __builtin_amdgcn_sched_barrier(0);
            static_for<kFB>([&](auto rr2) {
                constexpr int r2 = decltype(rr2)::value;
                if constexpr (r2 > r) fmac_row<r2 & 15>(x[r2], F[r & 1][r2 >> 4], x[r]);  // -= L11(r2, r) x_r
            });
        });
#pragma unroll
        for (int r = 0; r < kFB; ++r) col[r] = x[r];
    }
}

// C -= L21 U12 for one kFB x kFB tile, computed transposed on the matrix cores:
// C^T -= U12^T L21^T with v_mfma_f64_16x16x4f64 (lane l: A[l % 16][l / 16],
// B[l / 16][l % 16], D[4 e + l / 16][l % 16] for e = 0..3), so that C and the
// L21 operand are read down columns (coalesced) and only U12 goes through LDS.
// Wave w owns the 32 x 32 quadrant rows 32 (w >> 1), columns 32 (w & 1) of C.
// QuadUpdate is that quadrant's part for the pivoted factorization below;
// lu_update_kernel writes the same steps out with its own addressing.
struct QuadUpdate {
    f64x4 acc[2][2];  // [tc][tr]: C^T tile (c = cw + 16 tc + 4 e + lk, r = rw + 16 tr + lr)
    int lr, lk, rw, cw;
    __device__ QuadUpdate(int w, int l) : lr(l & 15), lk(l >> 4), rw(32 * (w >> 1)), cw(32 * (w & 1)) {}
    // the quadrant's entries of the tile at C (column-major, leading dimension lda)
    __device__ void load(const double* __restrict__ C, size_t lda)
    {
#pragma unroll
        for (int tc = 0; tc < 2; ++tc)
#pragma unroll
            for (int tr = 0; tr < 2; ++tr)
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[tc][tr][e] = C[(cw + 16 * tc + 4 * e + lk) * lda + rw + 16 * tr + lr];
    }
    // acc -= L21 U12: L21[r][k] at L[k * lda + r] (the tile's rows), U12[k][c] at U[k * kFLd + c] (LDS); the
    // L21 operands of eight k-steps loaded together (a 1024-thread workgroup has 128 VGPRs a lane)
    __device__ void apply(const double* __restrict__ L, size_t lda, const double* U)
    {
        const double* Lq = L + rw + lr;
#pragma unroll
        for (int h = 0; h < kFB / 4; h += 8) {
            double bl[8][2];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const size_t k = 4 * (h + s) + lk;
                bl[s][0] = Lq[k * lda];
                bl[s][1] = Lq[k * lda + 16];
            }
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int k = 4 * (h + s) + lk;
                const double a0 = U[k * kFLd + cw + lr], a1 = U[k * kFLd + cw + 16 + lr];  // U12^T[c][k]
                const double b0 = -bl[s][0], b1 = -bl[s][1];                                 // -L21^T[k][r]
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
            }
        }
    }
    __device__ void store(double* __restrict__ C, size_t lda) const
    {
#pragma unroll
        for (int tc = 0; tc < 2; ++tc)
#pragma unroll
            for (int tr = 0; tr < 2; ++tr)
#pragma unroll
                for (int e = 0; e < 4; ++e) C[(cw + 16 * tc + 4 * e + lk) * lda + rw + 16 * tr + lr] = acc[tc][tr][e];
    }
};

__global__ __launch_bounds__(256) void lu_update_kernel(const FacDesc* __restrict__ fd, const int* __restrict__ pre,
                                                        const int* __restrict__ tb, int count,
                                                        double* __restrict__ A, int j0)
{
    __shared__ double U[kFB * kFLd];  // U12 tile, U[k * kFLd + c]
    const int g = xcd_item();
    const int i = owner(pre, count, g);
    const FacDesc m = fd[i];
    const size_t lda = (size_t)m.npad;
    const int T = tb[i];
    const int item = g - pre[i];
    const int bi = item / T + 1, bj = item % T + 1;
    const size_t r0 = (size_t)j0 + kFB * bi, c0 = (size_t)j0 + kFB * bj;
    const double* Ab = A + m.a_off;
    const int t = threadIdx.x, lo = t & 63, q = t >> 6;
    for (int c = q; c < kFB; c += 4) U[lo * kFLd + c] = Ab[(c0 + c) * lda + j0 + lo];  // U12[k = lo][c]
    const int w = t >> 6, l = t & 63, lr = l & 15, lk = l >> 4;
    const int rw = 32 * (w >> 1), cw = 32 * (w & 1);
    double* Cb = A + m.a_off;
    f64x4 acc[2][2];  // [tc][tr]: C^T tile (c = cw + 16 tc + 4 e + lk, r = rw + 16 tr + lr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
        for (int tr = 0; tr < 2; ++tr)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                acc[tc][tr][e] = Cb[(c0 + cw + 16 * tc + 4 * e + lk) * lda + r0 + rw + 16 * tr + lr];
    __syncthreads();
    // the L21 operands of all 16 k-steps first (independent loads in flight together)
    const double* L = Ab + (size_t)j0 * lda + r0 + rw + lr;  // L21[r][k] at L[k * lda + 16 tr]
    double bl[kFB / 4][2];
#pragma unroll
    for (int s = 0; s < kFB / 4; ++s) {
        const size_t k = 4 * s + lk;
        bl[s][0] = L[k * lda];
        bl[s][1] = L[k * lda + 16];
    }
#pragma unroll
    for (int s = 0; s < kFB / 4; ++s) {
        const int k = 4 * s + lk;
        const double a0 = U[k * kFLd + cw + lr], a1 = U[k * kFLd + cw + 16 + lr];  // U12^T[c][k]
        const double b0 = -bl[s][0], b1 = -bl[s][1];                                 // -L21^T[k][r]
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
        for (int tr = 0; tr < 2; ++tr)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                Cb[(c0 + cw + 16 * tc + 4 * e + lk) * lda + r0 + rw + 16 * tr + lr] = acc[tc][tr][e];
}

// a valid CSC matrix: Ap[0] = 0, nondecreasing column pointers, rows in [0, n)
bool valid_csc(const BatchDenseLU::Matrix& m)
{
    if (m.Ap[0] != 0) return false;
    for (int c = 0; c < m.n; ++c) {
        if (m.Ap[c + 1] < m.Ap[c]) return false;
        for (int k = m.Ap[c]; k < m.Ap[c + 1]; ++k)
            if (m.Ai[k] < 0 || m.Ai[k] >= m.n) return false;
    }
    return true;
}

// The launch plan of the blocked no-pivot LU over matrices fd (largest first): per step p the number of
// active matrices (a prefix), and offsets into pre of their trailing block counts T (tbo), of the panel
// items' prefix sums (pre2, n2 items, padded to fours per matrix) and of the update tiles' prefix sums
// (pre3, n3 tiles).  last[i]: matrix i's envelope (its last nonzero block per step) or null (dense).
struct LuStep {
    int act, tbo, pre2, n2, pre3, n3;
};

bool plan_lu_steps(const std::vector<FacDesc>& fd, const std::vector<const int*>& last, std::vector<LuStep>& steps,
                   std::vector<int>& pre)
{
    const int Ka = (int)fd.size();
    const int P = Ka ? fd[0].npad / kFB : 0;
    steps.assign(P, LuStep{});
    pre.clear();
    std::vector<int> T(Ka);
    for (int p = 0; p < P; ++p) {
        int act = 0;
        while (act < Ka && fd[act].npad > kFB * p) ++act;
        LuStep& sp = steps[p];
        sp.act = act;
        sp.tbo = (int)pre.size();
        for (int i = 0; i < act; ++i) {
            T[i] = last[i] ? std::max(0, last[i][p] - p) : fd[i].npad / kFB - p - 1;
            pre.push_back(T[i]);
        }
        sp.pre2 = (int)pre.size();
        pre.push_back(0);
        for (int i = 0; i < act; ++i) pre.push_back(pre.back() + (2 * T[i] + 3) / 4 * 4);
        sp.n2 = pre.back();
        sp.pre3 = (int)pre.size();
        pre.push_back(0);
        for (int i = 0; i < act; ++i) {
            const long long t2 = (long long)T[i] * T[i];
            if (pre.back() + t2 > INT_MAX) return false;
            pre.push_back(pre.back() + (int)t2);
        }
        sp.n3 = pre.back();
    }
    return true;
}

int launch_lu_steps(const std::vector<LuStep>& steps, const FacDesc* fdesc, const int* pre, double* A, hipStream_t st)
{
    for (size_t p = 0; p < steps.size(); ++p) {
        const LuStep& sp = steps[p];
        const int j0 = kFB * (int)p;
        hipLaunchKernelGGL(lu_diag_kernel, dim3(sp.act), dim3(kFB), 0, st, fdesc, A, j0);
        if (sp.n2 > 0)
            hipLaunchKernelGGL(lu_panel_kernel, dim3(sp.n2 / 4), dim3(256), 0, st, fdesc, pre + sp.pre2,
                               pre + sp.tbo, sp.act, A, j0);
        if (sp.n3 > 0)
            hipLaunchKernelGGL(lu_update_kernel, dim3(sp.n3), dim3(256), 0, st, fdesc, pre + sp.pre3, pre + sp.tbo,
                               sp.act, A, j0);
    }
    return hipGetLastError() != hipSuccess ? -1 : 0;
}

// Partial pivoting LU of each matrix whose flag is set, one workgroup (16 wavefronts) per matrix, in place
// in its store, blocked like LAPACK's getrf by kFB columns: the panel (columns j0 .. j0 + kFB - 1, rows j0 ..
// n - 1) factored column by column as getf2 does (the first row of largest magnitude in the column, the
// panel's rows swapped, multipliers by division, the panel's later columns updated), the panel's row
// interchanges applied to the other columns in order, U12 = L11^-1 A12 (one wavefront per column of A12, the
// unit lower L11 in LDS), and A22 -= L21 U12 on the matrix cores, four 64 x 64 tiles at a time (QuadUpdate,
// U12 tiles through LDS).  perm (npad entries per matrix, at its solve descriptor's perm_off) holds the
// pivots during the factorization and then the solve's row order: the identity for the matrices whose flag
// is clear.  singular[k]: whether a zero pivot was met (its column is then left as it is, as getf2 leaves
// it).  The rare fallback of a factorization whose multipliers exceed the threshold: launches only, so it
// can sit in a captured graph, with one compute unit per flagged matrix.
// Rows and columns of the padding (n .. npad - 1) hold the identity and no interchange reaches them, so
// the trailing update runs over whole tiles up to npad: their L21 rows and U12 columns are zero.
constexpr int kPivThreads = 1024;
constexpr int kPivTiles = kPivThreads / 256;  // trailing-update tiles in flight, four wavefronts each

__global__ __launch_bounds__(kPivThreads) void lu_pivoted_kernel(const MatDesc* __restrict__ desc,
                                                                 const SolveDesc* __restrict__ sdesc,
                                                                 double* __restrict__ A, const int* __restrict__ flag,
                                                                 int* __restrict__ perm, int* __restrict__ singular)
{
    // U12 tiles of the update; L11 of the row-block solve (the first); the permutation at the end (as ints)
    __shared__ double tb[kPivTiles * kFB * kFLd];
    __shared__ double rv[kPivThreads / 64];
    __shared__ int ri[kPivThreads / 64];
    __shared__ int pk;
    const MatDesc m = desc[blockIdx.x];
    const int n = m.n, npad = m.npad, tid = threadIdx.x;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    int* pm = perm + sdesc[blockIdx.x].perm_off;
    if (!flag[blockIdx.x]) {
        for (int i = tid; i < npad; i += kPivThreads) pm[i] = i;
        if (tid == 0) singular[blockIdx.x] = 0;
        return;
    }
    double* a = A + m.a_off;
    const size_t ld = (size_t)npad;
    int sing = 0;
    for (int j0 = 0; j0 < n; j0 += kFB) {
        const int je = min(j0 + kFB, n);
        // the panel, column by column
        for (int k = j0; k < je; ++k) {
            double* ck = a + (size_t)k * ld;
            // the pivot: largest |a_rk|, r >= k, the first one on ties
            double bv = -1.0;
            int bi = n;
            for (int r = k + tid; r < n; r += kPivThreads) {
                const double v = std::fabs(ck[r]);
                if (v > bv) { bv = v; bi = r; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_down(bv, o);
                const int oi = __shfl_down(bi, o);
                if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
            }
            if (lane == 0) { rv[wv] = bv; ri[wv] = bi; }
            __syncthreads();
            if (tid == 0) {
                for (int w = 1; w < kPivThreads / 64; ++w)
                    if (rv[w] > rv[0] || (rv[w] == rv[0] && ri[w] < ri[0])) { rv[0] = rv[w]; ri[0] = ri[w]; }
                pk = ri[0] < n ? ri[0] : k;
                pm[k] = pk;
            }
            __syncthreads();
            const int p = pk;
            if (p != k && tid < je - j0) {  // the panel's two rows
                double* c = a + (size_t)(j0 + tid) * ld;
                const double t = c[k];
                c[k] = c[p];
                c[p] = t;
            }
            __syncthreads();
            const double d = ck[k];
            if (d == 0.0) {
                sing = 1;
                continue;  // uniform: every thread read the same pivot
            }
            for (int r = k + 1 + tid; r < n; r += kPivThreads) ck[r] /= d;
            __syncthreads();
            // the panel's later columns, rows below k: a_rc -= a_rk a_kc
            for (int r = k + 1 + tid; r < n; r += kPivThreads) {
                const double l = ck[r];
                for (int c = k + 1; c < je; ++c) {
                    double* cc = a + (size_t)c * ld;
                    cc[r] -= l * cc[k];
                }
            }
            __syncthreads();
        }
        // the panel's interchanges on the columns outside it, in pivot order
        for (int c = tid; c < n; c += kPivThreads) {
            if (c >= j0 && c < je) continue;
            double* cc = a + (size_t)c * ld;
            for (int k = j0; k < je; ++k) {
                const int p = pm[k];
                if (p != k) {
                    const double t = cc[k];
                    cc[k] = cc[p];
                    cc[p] = t;
                }
            }
        }
        if (je >= n) break;  // the last panel: no columns right of it
        // U12 = L11^-1 A12: L11 (unit lower, rows and columns j0 .. j0 + kFB - 1) into LDS as L[k][i]
        double* L11 = tb;
        for (int e = tid; e < kFB * kFB; e += kPivThreads) {
            const int i = e % kFB, k = e / kFB;
            L11[k * kFLd + i] = a[(size_t)(j0 + k) * ld + j0 + i];
        }
        __syncthreads();
        for (int c = je + wv; c < n; c += kPivThreads / 64) {
            double* cc = a + (size_t)c * ld + j0;
            double x = cc[lane];
            for (int k = 0; k < kFB - 1; ++k) {
                const double xk = lane_value(x, k);
                if (lane > k) x = fma(-L11[k * kFLd + lane], xk, x);
            }
            cc[lane] = x;
        }
        __syncthreads();
        // A22 -= L21 U12 over the tiles of rows and columns je .. npad - 1 (je = j0 + kFB here)
        const int T = (npad - je) / kFB;
        const int tg = wv >> 2;
        double* U = tb + tg * kFB * kFLd;
        for (int t0 = 0; t0 < T * T; t0 += kPivTiles) {
            const int t = t0 + tg;
            const bool act = t < T * T;
            const size_t r0 = (size_t)je + kFB * (act ? t / T : 0), c0 = (size_t)je + kFB * (act ? t % T : 0);
            QuadUpdate u(wv & 3, lane);
            if (act) {
                for (int c = wv & 3; c < kFB; c += 4) U[lane * kFLd + c] = a[(c0 + c) * ld + j0 + lane];  // U12[k][c]
                u.load(a + c0 * ld + r0, ld);
            }
            __syncthreads();
            if (act) {
                u.apply(a + (size_t)j0 * ld + r0, ld, U);
                u.store(a + c0 * ld + r0, ld);
            }
            __syncthreads();
        }
    }
    // the pivots into the permutation (rows beyond n stay in place)
    int* q = reinterpret_cast<int*>(tb);
    __syncthreads();
    for (int i = tid; i < npad; i += kPivThreads) q[i] = i;
    __syncthreads();
    if (tid == 0)
        for (int i = 0; i < n; ++i) {
            const int j = pm[i];
            if (j != i) {
                const int t = q[i];
                q[i] = q[j];
                q[j] = t;
            }
        }
    __syncthreads();
    for (int i = tid; i < npad; i += kPivThreads) pm[i] = q[i];
    if (tid == 0) singular[blockIdx.x] = sing;
}

class BatchLU : public BatchDenseLU {
public:
    explicit BatchLU(int device) : device_(device)
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 65536;
        lds_bytes_ = lds;
        lds_order_ = std::max(0, (lds / 8 - kSolveLds) / 256 * 256);
    }

    ~BatchLU() override
    {
        for (rocblas_handle h : side_h_) (void)rocblas_destroy_handle(h);
        for (hipStream_t st : side_) (void)hipStreamDestroy(st);
        for (hipEvent_t e : ev_)
            if (e) (void)hipEventDestroy(e);
    }

    void set_parallel(ParallelFor pf, void* user) override
    {
        pf_ = pf;
        user_ = user;
    }

    int factor(int K, const Matrix* m, unsigned char* singular, unsigned char* unpivoted) override
    {
        dense_ = DenseSrc();
        return factor_impl(K, m, singular, unpivoted);
    }

    int factor_dense_dev(int n, const double* a, long long rs, long long cs, unsigned char* singular,
                         unsigned char* unpivoted) override
    {
        K_ = 0;
        if (n <= 0 || n > kMaxOrder || !a) return -1;
        const Matrix m{n, nullptr, nullptr, nullptr};
        dense_ = DenseSrc{a, rs, cs};
        const int rc = factor_impl(1, &m, singular, unpivoted);
        dense_ = DenseSrc();
        return rc;
    }

    int solve_dense_dev(int n, const double* b, double* x) override
    {
        if (K_ != 1 || n != (int)b_off_[1] || !b || !x) return -1;
        ThreadRes* r = thread_res(device_);
        if (!r || hipSetDevice(device_) != hipSuccess) return -2;
        hipStream_t st = r->stream;
        if (hipMemcpyAsync(b_.p, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess) return -2;
        if (launch_solve(n, st) || hipMemcpyAsync(x, b_.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return -2;
        return 0;
    }

    int factor_impl(int K, const Matrix* m, unsigned char* singular, unsigned char* unpivoted)
    {
        K_ = 0;
        if (K < 0 || (K > 0 && !m)) return -1;
        if (K == 0) return 0;
        if (K > kMaxMatrices) return -1;
        const bool dense = dense_.a != nullptr;
        for (int k = 0; k < K; ++k)
            if (m[k].n < 0 || (!dense && m[k].n > 0 && (!m[k].Ap || !m[k].Ai || !m[k].Ax)) || m[k].n > kMaxOrder ||
                (m[k].dAp && (!m[k].dAi || !m[k].dAx)))
                return -1;
        std::vector<unsigned char> ok(K, 1);
        if (!dense) par(K, [&](int k) { ok[k] = m[k].n == 0 || valid_csc(m[k]); });
        for (int k = 0; k < K; ++k)
            if (!ok[k]) return -1;
        ThreadRes* r = thread_res(device_);
        if (!r || hipSetDevice(device_) != hipSuccess) return -2;
        hipStream_t st = r->stream;

        // layout in the caller's order
        desc_h_.resize(K);
        std::vector<long long> nz_off(K + 1, 0), col_off(K, 0);
        b_off_.assign(K + 1, 0);
        size_t a_total = 0;
        long long col_total = 0, piv_total = 0;
        int npad_max = 0;
        for (int k = 0; k < K; ++k) {
            const int n = m[k].n, npad = gpu_lu_padded(n);
            MatDesc& d = desc_h_[k];
            d.a_off = (long long)a_total;
            d.n = n;
            d.npad = npad;
            d.ipiv_off = -1;
            d.ord_off = d.env_off = -1;
            d.row_base = 0;
            d.ap = d.ai = nullptr;
            d.ax = nullptr;
            col_off[k] = col_total;
            a_total += (size_t)npad * npad;
            if (!dense && !m[k].dAp) col_total += n + 1;  // staged from the host
            piv_total += npad;
            nz_off[k + 1] = nz_off[k] + (n > 0 && !dense && !m[k].dAp ? m[k].Ap[n] : 0);
            b_off_[k + 1] = b_off_[k] + n;
            npad_max = std::max(npad_max, npad);
        }
        if (nz_off[K] > INT_MAX || col_total > INT_MAX || piv_total > INT_MAX / 2) return -1;
        const long long nnz = nz_off[K];

        // each matrix's symmetric order and block envelope, from its pattern
        // (a dense device matrix keeps its order and is factored whole)
        plan_.resize(K);
        cur_.assign(K, nullptr);
        {
            std::set<int> seen;  // a key named twice in one call: those matrices keep slot plans
            for (int k = 0; k < K; ++k)
                cur_[k] = m[k].key >= 0 && seen.insert(m[k].key).second ? &keyed_[m[k].key] : &plan_[k];
        }
        long long aux_total = 0;
        if (!dense) {
            par(K, [&](int k) {
                Plan& pl = *cur_[k];
                const int n = m[k].n;
                // the order depends on the pattern alone: kept while slot k's pattern repeats
                if (n > 0 && pl.n == n && std::equal(m[k].Ap, m[k].Ap + n + 1, pl.pat_p.begin()) &&
                    pl.pat_i.size() == (size_t)m[k].Ap[n] && std::equal(m[k].Ai, m[k].Ai + m[k].Ap[n], pl.pat_i.begin()))
                    return;
                // a keyed matrix whose pattern changed keeps its order while that order's envelope of the new
                // pattern stays within 1/8 of the envelope it was chosen with
                const bool keep = cur_[k] != &plan_[k] && n > 0 && pl.n == n && pl.base_tiles > 0;
                pl.n = n;
                pl.pat_p.assign(m[k].Ap, m[k].Ap + (n > 0 ? n + 1 : 0));
                pl.pat_i.assign(m[k].Ai, m[k].Ai + (n > 0 ? m[k].Ap[n] : 0));
                if (keep) {
                    BlockEnvelope e = block_envelope(n, desc_h_[k].npad, kFB, m[k].Ap, m[k].Ai,
                                                     pl.re ? pl.inv.data() : nullptr, m[k].symmetric);
                    if (8 * e.tiles <= 9 * pl.base_tiles) {
                        pl.env = std::move(e);
                        return;
                    }
                }
                pl.re = n > 0 && band_order(n, desc_h_[k].npad, kFB, m[k].Ap, m[k].Ai, pl.ord, pl.inv, pl.env,
                                            m[k].symmetric);
                pl.base_tiles = pl.env.tiles;
            });
            for (int k = 0; k < K; ++k) {
                MatDesc& d = desc_h_[k];
                if (d.n == 0) continue;
                if (cur_[k]->re) {
                    d.ord_off = (int)aux_total;
                    aux_total += 2LL * d.n;
                }
                d.env_off = (int)aux_total;
                aux_total += 3LL * (d.npad / kFB);
            }
            if (aux_total > INT_MAX) return -1;
            if (!aux_.ensure((size_t)aux_total) || !aux_h_.ensure((size_t)aux_total)) return -2;
            par(K, [&](int k) {
                const MatDesc& d = desc_h_[k];
                const Plan& pl = *cur_[k];
                if (d.ord_off >= 0) {
                    std::copy(pl.ord.begin(), pl.ord.end(), aux_h_.p + d.ord_off);
                    std::copy(pl.inv.begin(), pl.inv.end(), aux_h_.p + d.ord_off + d.n);
                }
                if (d.env_off >= 0) {
                    const int nb = d.npad / kFB;
                    std::copy(pl.env.last.begin(), pl.env.last.end(), aux_h_.p + d.env_off);
                    std::copy(pl.env.top.begin(), pl.env.top.end(), aux_h_.p + d.env_off + nb);
                    std::copy(pl.env.first.begin(), pl.env.first.end(), aux_h_.p + d.env_off + 2 * nb);
                }
            });
        }

        // every matrix first without pivoting (largest first: that is also the
        // solve's launch order)
        std::vector<int> order = by_order(K, [](int k) { return k; });
        std::vector<int> pos(K);
        for (int i = 0; i < K; ++i) pos[order[i]] = i;

        if (!A_.ensure(a_total) || !ipiv_.ensure((size_t)piv_total) || !iscr_.ensure(2 * (size_t)piv_total) ||
            !info_.ensure(3 * (size_t)K) || !desc_.ensure(2 * (size_t)K) ||
            !sdesc_.ensure(K) || !colp_.ensure((size_t)col_total) || !rowi_.ensure((size_t)nnz) ||
            !vals_.ensure((size_t)nnz) || !b_.ensure(2 * (size_t)b_off_[K]))
            return -2;
        if (!colp_h_.ensure((size_t)col_total) || !rowi_h_.ensure((size_t)nnz) || !vals_h_.ensure((size_t)nnz) ||
            !info_h_.ensure(3 * (size_t)K))
            return -2;
        int* info = info_.p;          // [0, K): no-pivot pass (group order)
        int* flag = info_.p + K;      // [K, 2K): check (caller's order)
        int* info2 = info_.p + 2 * K; // [2K, 3K): pivoted pass (its group order)

        if (!dense) par(K, [&](int k) {
            const int n = m[k].n;
            MatDesc& d = desc_h_[k];
            if (m[k].dAp) {  // already on the device
                d.ap = m[k].dAp;
                d.ai = m[k].dAi;
                d.ax = m[k].dAx;
                d.row_base = m[k].row_base;
                return;
            }
            d.ap = colp_.p + col_off[k];
            d.ai = rowi_.p;
            d.ax = vals_.p;
            if (n == 0) return;
            const long long base = nz_off[k];
            int* cp = colp_h_.p + col_off[k];
            for (int c = 0; c <= n; ++c) cp[c] = (int)(base + m[k].Ap[c]);
            const int cnt = m[k].Ap[n];
            std::copy(m[k].Ai, m[k].Ai + cnt, rowi_h_.p + base);
            std::copy(m[k].Ax, m[k].Ax + cnt, vals_h_.p + base);
        });
        if (hipMemcpyAsync(desc_.p, desc_h_.data(), sizeof(MatDesc) * K, hipMemcpyHostToDevice, st) != hipSuccess ||
            (aux_total > 0 &&
             hipMemcpyAsync(aux_.p, aux_h_.p, sizeof(int) * aux_total, hipMemcpyHostToDevice, st) != hipSuccess) ||
            (col_total > 0 &&
             hipMemcpyAsync(colp_.p, colp_h_.p, sizeof(int) * col_total, hipMemcpyHostToDevice, st) != hipSuccess) ||
            (nnz > 0 &&
             (hipMemcpyAsync(rowi_.p, rowi_h_.p, sizeof(int) * nnz, hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(vals_.p, vals_h_.p, sizeof(double) * nnz, hipMemcpyHostToDevice, st) != hipSuccess)))
            return -2;
        if (assemble(desc_.p, K, npad_max, st)) return -2;
        if (hipMemsetAsync(info, 0, sizeof(int) * 2 * K, st) != hipSuccess) return -2;
        if (factor_unpivoted(order, st)) return -2;
        for (int k0 = 0; k0 < K; k0 += kGridY) {
            const int kc = std::min(kGridY, K - k0);
            if (npad_max > 0)
            {
                const int sp = column_split((long long)(npad_max / kFB) * kc);
                hipLaunchKernelGGL(multiplier_check_kernel, dim3(npad_max / kFB * sp, kc), dim3(256), 0, st,
                                   desc_.p + k0, A_.p, flag + k0, aux_.p, sp);
            }
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(info_h_.p, info, sizeof(int) * 2 * K, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return -2;

        // the rest again from the CSC, with partial pivoting
        std::vector<int> redo;
        for (int k = 0; k < K; ++k)
            if (desc_h_[k].n > 0 && (info_h_.p[pos[k]] > 0 || info_h_.p[K + k] != 0)) redo.push_back(k);
        const int F = (int)redo.size();
        std::vector<int> redo_pos(K, -1);
        if (F > 0) {
            std::vector<int> order2 = by_order(F, [&](int i) { return redo[i]; });
            for (int& i : order2) i = redo[i];
            // one rocSOLVER getrf per matrix (its blocked panels: ~34 ms at order 4096, where getrf_batched
            // with one matrix steps column by column and took ~180 ms); a matrix gets the same bits from
            // it whatever else is refactored in the call
            std::vector<Group> groups2;
            for (int i = 0; i < F; ++i) groups2.push_back({i, 1, desc_h_[order2[i]].npad});
            std::vector<MatDesc> sub(F);
            std::vector<double*> redo_a(F);  // each matrix's store
            int npad2 = 0;
            long long poff = 0;
            for (int i = 0; i < F; ++i) {
                const int k = order2[i];
                desc_h_[k].ipiv_off = (int)poff;
                poff += desc_h_[k].npad;
                sub[i] = desc_h_[k];
                sub[i].env_off = -1;  // assembled and factored whole (in the same order)
                redo_pos[k] = i;
                redo_a[i] = A_.p + desc_h_[k].a_off;
                npad2 = std::max(npad2, desc_h_[k].npad);
            }
            if (hipMemcpyAsync(desc_.p + K, sub.data(), sizeof(MatDesc) * F, hipMemcpyHostToDevice, st) != hipSuccess)
                return -2;
            if (assemble(desc_.p + K, F, npad2, st)) return -2;
            if (hipMemsetAsync(info2, 0, sizeof(int) * F, st) != hipSuccess) return -2;
            if (run_groups(groups2, r, [&](rocblas_handle h, const Group& g) {
                    return rocsolver_dgetrf(h, g.npad, g.npad, redo_a[g.first], g.npad,
                                            ipiv_.p + desc_h_[order2[g.first]].ipiv_off, info2 + g.first);
                }))
                return -2;
            if (npad2 <= lds_bytes_ / 8)
                hipLaunchKernelGGL(pivots_to_perm_kernel<true>, dim3(F), dim3(256), sizeof(int) * 2 * npad2, st,
                                   desc_.p + K, ipiv_.p, iscr_.p);
            else
                hipLaunchKernelGGL(pivots_to_perm_kernel<false>, dim3(F), dim3(256), 0, st, desc_.p + K, ipiv_.p,
                                   iscr_.p);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(info_h_.p + 2 * K, info2, sizeof(int) * F, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return -2;
        }
        for (int k = 0; k < K; ++k) {
            const bool pivoted = redo_pos[k] >= 0;
            // U(info, info) == 0 after partial pivoting: UMFPACK_WARNING_singular_matrix's case
            if (singular) singular[k] = pivoted && info_h_.p[2 * K + redo_pos[k]] > 0;
            if (unpivoted) unpivoted[k] = !pivoted;
        }

        // the solves' descriptors, largest matrices first
        std::vector<SolveDesc> sd;
        sd.reserve(K);
        for (int i = 0; i < K; ++i) {
            const int k = order[i];
            if (desc_h_[k].n == 0) continue;
            const bool pivoted = redo_pos[k] >= 0;
            sd.push_back({desc_h_[k].a_off, desc_h_[k].n, desc_h_[k].npad, (int)b_off_[k],
                          pivoted ? desc_h_[k].ipiv_off : -1, desc_h_[k].ord_off, pivoted ? -1 : desc_h_[k].env_off});
        }
        nsolve_ = (int)sd.size();
        npad_max_ = npad_max;
        if (nsolve_ > 0 &&
            (hipMemcpyAsync(sdesc_.p, sd.data(), sizeof(SolveDesc) * nsolve_, hipMemcpyHostToDevice, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess))
            return -2;
        ndesc_.resize(K);
        for (int k = 0; k < K; ++k) ndesc_[k] = desc_h_[k].n;
        K_ = K;
        return 0;
    }

    int solve(int K, const double* const* b, double* const* x) override
    {
        if (K != K_ || (K > 0 && (!b || !x))) return -1;
        if (K == 0 || nsolve_ == 0) return 0;
        ThreadRes* r = thread_res(device_);
        if (!r || hipSetDevice(device_) != hipSuccess) return -2;
        const long long total = b_off_[K];
        if (!bh_.ensure((size_t)total)) return -2;
        par(K, [&](int k) { std::copy(b[k], b[k] + ndesc_[k], bh_.p + b_off_[k]); });
        hipStream_t st = r->stream;
        if (hipMemcpyAsync(b_.p, bh_.p, sizeof(double) * total, hipMemcpyHostToDevice, st) != hipSuccess) return -2;
        if (launch_solve(total, st) ||
            hipMemcpyAsync(bh_.p, b_.p, sizeof(double) * total, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return -2;
        par(K, [&](int k) { std::copy(bh_.p + b_off_[k], bh_.p + b_off_[k] + ndesc_[k], x[k]); });
        return 0;
    }

private:
    static constexpr int kGridY = 65535;  // grid.y limit: matrices per launch
    static constexpr int kMaxMatrices = 1 << 24;
    static constexpr int kMaxOrder = 46080;  // npad^2 stays below 2^31 (rocSOLVER's int lda * n)

    struct Group {
        int first, count, npad;
    };

    static constexpr int kSide = 3;  // side streams for the group calls of one factorization

    int device_;
    int lds_bytes_ = 65536, lds_order_ = 0;
    std::vector<hipStream_t> side_;
    std::vector<rocblas_handle> side_h_;
    std::vector<hipEvent_t> ev_;  // [0]: inputs ready on the caller's stream, [1 + s]: side stream s done
    bool side_failed_ = false;
    ParallelFor pf_ = nullptr;
    void* user_ = nullptr;
    int K_ = 0, nsolve_ = 0, npad_max_ = 0;
    std::vector<MatDesc> desc_h_;
    std::vector<long long> b_off_;
    std::vector<int> ndesc_;
    DevBuf<double> A_, vals_, b_;
    DevBuf<int> ipiv_, iscr_, info_, colp_, rowi_;
    struct DenseSrc {
        const double* a = nullptr;
        long long rs = 0, cs = 0;
    };
    DenseSrc dense_;  // factor_dense_dev's source while it runs (assemble() copies from it)
    DevBuf<MatDesc> desc_;
    struct Plan {
        bool re = false;  // reordered: ord (ord[new] = old) and inv hold the order
        std::vector<int> ord, inv;
        BlockEnvelope env;
        int n = -1;  // the pattern the order was taken from
        std::vector<int> pat_p, pat_i;
        long long base_tiles = 0;  // the envelope's tiles when the order was chosen
    };
    std::vector<Plan> plan_;      // per matrix slot of the last factor() (CSC input)
    std::map<int, Plan> keyed_;   // per caller key (Matrix::key >= 0)
    std::vector<Plan*> cur_;      // the plan of each matrix of the last factor()
    DevBuf<int> aux_;         // orders and envelopes, at MatDesc::ord_off / env_off
    PinBuf<int> aux_h_;
    DevBuf<FacDesc> fdesc_;
    DevBuf<int> fpre_;
    DevBuf<SolveDesc> sdesc_;
    PinBuf<int> colp_h_, rowi_h_, info_h_;
    PinBuf<double> vals_h_, bh_;

    // indices 0..count-1 (through key to the caller's index) by padded order, largest first, stable
    template <class Key>
    std::vector<int> by_order(int count, Key key) const
    {
        std::vector<int> o(count);
        for (int i = 0; i < count; ++i) o[i] = i;
        std::stable_sort(o.begin(), o.end(),
                         [&](int a, int b) { return desc_h_[key(a)].npad > desc_h_[key(b)].npad; });
        return o;
    }

    // the blocked no-pivot LU of every matrix, in `order` (largest first):
    // three launches per step of kFB columns, each over all matrices still active
    int factor_unpivoted(const std::vector<int>& order, hipStream_t st)
    {
        std::vector<FacDesc> fd;
        std::vector<const int*> last;  // each matrix's envelope (last block per step), or null
        for (int k : order)
            if (desc_h_[k].npad > 0) {
                fd.push_back({desc_h_[k].a_off, desc_h_[k].npad, 0});
                last.push_back(desc_h_[k].env_off >= 0 ? cur_[k]->env.last.data() : nullptr);
            }
        if (fd.empty()) return 0;
        std::vector<LuStep> steps;
        std::vector<int> pre;
        if (!plan_lu_steps(fd, last, steps, pre)) return -1;
        if (!fdesc_.ensure(fd.size()) || !fpre_.ensure(pre.size())) return -1;
        if (hipMemcpyAsync(fdesc_.p, fd.data(), sizeof(FacDesc) * fd.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(fpre_.p, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice, st) != hipSuccess)
            return -1;
        return launch_lu_steps(steps, fdesc_.p, fpre_.p, A_.p, st);
    }

    bool ensure_side()
    {
        if (!side_.empty()) return true;
        if (side_failed_) return false;
        std::vector<hipStream_t> st(kSide, nullptr);
        std::vector<rocblas_handle> h(kSide, nullptr);
        std::vector<hipEvent_t> ev(kSide + 1, nullptr);
        bool ok = hipSetDevice(device_) == hipSuccess;
        for (int i = 0; ok && i < kSide; ++i)
            ok = hipStreamCreateWithFlags(&st[i], hipStreamNonBlocking) == hipSuccess &&
                 rocblas_create_handle(&h[i]) == rocblas_status_success &&
                 rocblas_set_stream(h[i], st[i]) == rocblas_status_success;
        for (int i = 0; ok && i <= kSide; ++i) ok = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            for (rocblas_handle x : h)
                if (x) (void)rocblas_destroy_handle(x);
            for (hipStream_t x : st)
                if (x) (void)hipStreamDestroy(x);
            for (hipEvent_t x : ev)
                if (x) (void)hipEventDestroy(x);
            side_failed_ = true;
            return false;
        }
        side_ = st;
        side_h_ = h;
        ev_ = ev;
        return true;
    }

    // call(handle, group) for every group: one group runs on the caller's
    // stream; several are dealt round-robin (largest first) over that stream
    // and kSide side streams, which start after the work already queued on the
    // caller's stream and finish before anything queued on it afterwards.
    // Each matrix goes through the same call whichever stream runs it.
    template <class Call>
    int run_groups(const std::vector<Group>& groups, ThreadRes* r, Call call)
    {
        if (groups.size() <= 1 || !ensure_side()) {
            for (const Group& g : groups)
                if (call(r->handle, g) != rocblas_status_success) return -2;
            return 0;
        }
        if (hipEventRecord(ev_[0], r->stream) != hipSuccess) return -2;
        for (int i = 0; i < kSide; ++i)
            if (hipStreamWaitEvent(side_[i], ev_[0], 0) != hipSuccess) return -2;
        for (size_t i = 0; i < groups.size(); ++i) {
            const int lane = (int)(i % (kSide + 1));
            if (call(lane == 0 ? r->handle : side_h_[lane - 1], groups[i]) != rocblas_status_success) return -2;
        }
        for (int i = 0; i < kSide; ++i)
            if (hipEventRecord(ev_[1 + i], side_[i]) != hipSuccess ||
                hipStreamWaitEvent(r->stream, ev_[1 + i], 0) != hipSuccess)
                return -2;
        return 0;
    }

    // the solve kernel over the right-hand sides packed in b_ (total entries; x in global scratch after them
    // when the matrices are too large for LDS)
    int launch_solve(long long total, hipStream_t st)
    {
        if (nsolve_ == 0) return 0;
        if (npad_max_ <= lds_order_)
            hipLaunchKernelGGL(lu_solve_kernel<true>, dim3(nsolve_), dim3(kSolveThreads),
                               sizeof(double) * (kSolveLds + npad_max_), st, sdesc_.p, A_.p, ipiv_.p, b_.p,
                               b_.p + total, aux_.p, nullptr);
        else
            hipLaunchKernelGGL(lu_solve_kernel<false>, dim3(nsolve_), dim3(kSolveThreads),
                               sizeof(double) * kSolveLds, st, sdesc_.p, A_.p, ipiv_.p, b_.p, b_.p + total,
                               aux_.p, nullptr);
        return hipGetLastError() != hipSuccess ? -1 : 0;
    }

    int assemble(const MatDesc* desc, int K, int npad_max, hipStream_t st)
    {
        if (npad_max == 0) return 0;
        if (dense_.a) {  // one matrix from the device
            hipLaunchKernelGGL(dense_copy_kernel, dim3(npad_max, 1), dim3(256), 0, st, desc, A_.p, dense_.a, dense_.rs,
                               dense_.cs);
            return hipGetLastError() != hipSuccess ? -1 : 0;
        }
        for (int k0 = 0; k0 < K; k0 += kGridY) {
            const int kc = std::min(kGridY, K - k0);
            const int sp = column_split((long long)(npad_max / kFB) * kc);
            hipLaunchKernelGGL(fill_identity_kernel, dim3(npad_max / kFB * sp, kc), dim3(256), 0, st, desc + k0, A_.p,
                               aux_.p, sp);
            hipLaunchKernelGGL(scatter_kernel, dim3((npad_max + 255) / 256, kc), dim3(256), 0, st, desc + k0, A_.p,
                               aux_.p);
        }
        return hipGetLastError() != hipSuccess ? -1 : 0;
    }

    template <class F>
    void par(int count, F&& f)
    {
        using Fn = std::remove_reference_t<F>;
        if (!pf_ || count < 2) {
            for (int k = 0; k < count; ++k) f(k);
            return;
        }
        pf_(user_, count, [](void* a, int k) { (*static_cast<Fn*>(a))(k); }, &f);
    }
};

class GpuDenseLU : public LinearSolver {
public:
    explicit GpuDenseLU(int device) : lu_(device) {}

    int factor(int n, const int* Ap, const int* Ai, const double* Ax) override
    {
        singular_ = unpivoted_ = false;
        if (n < 0 || (n > 0 && !Ap)) return -1;
        BatchDenseLU::Matrix m{n, Ap, Ai, Ax};
        m.key = 0;  // one simulation's matrices: its order follows its own sequence of patterns
        m.symmetric = true;  // the simulation's step matrices (V is stored iff V^T is)
        unsigned char s = 0, u = 0;
        const int rc = lu_.factor(1, &m, &s, &u);
        if (rc < 0) return rc;
        singular_ = s != 0;
        unpivoted_ = u != 0;
        return singular_ ? 1 : 0;
    }

    int solve(const double* b, double* x) override { return lu_.solve(1, &b, &x); }
    bool singular() const override { return singular_; }
    bool unpivoted() const override { return unpivoted_; }

private:
    BatchLU lu_;
    bool singular_ = false, unpivoted_ = false;
};

// The dense LU of a fixed batch of matrices by launches alone (see DenseLUGraph in d4_solver.h): the
// plans, descriptors and device arrays are made once here, so factor and solve allocate nothing, copy
// nothing from the host and never wait for the device.
class GraphLU : public DenseLUGraph {
public:
    GraphLU() = default;
    ~GraphLU() override
    {
        for (void* p : {(void*)fd_d_, (void*)pre_d_, (void*)md_d_, (void*)sd_d_, (void*)perm_d_, (void*)scratch_d_,
                        (void*)aux_d_})
            if (p) (void)hipFree(p);
    }

    int init(int device, int K, const int* n, const long long* b_off, const std::vector<std::vector<int>>* ord,
             const std::vector<BlockEnvelope>* env)
    {
        if (K <= 0) return -1;
        if (hipSetDevice(device) != hipSuccess) return -2;
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) lds = 65536;
        K_ = K;
        std::vector<MatDesc> md(K);
        std::vector<SolveDesc> sd(K);
        a_off_.resize(K);
        ord_off_.assign(K, -1);
        env_off_.assign(K, -1);
        std::vector<int> aux;
        long long a = 0, poff = 0, btot = 0;
        for (int k = 0; k < K; ++k) {
            const int npad = gpu_lu_padded(n[k]);
            if (n[k] <= 0 || npad * 8 > lds) return -1;  // the fallback keeps 2 npad ints in LDS
            MatDesc& m = md[k];
            m = MatDesc{};
            m.a_off = a_off_[k] = a;
            m.n = n[k];
            m.npad = npad;
            m.ipiv_off = -1;
            if (ord && !(*ord)[k].empty()) {  // ord then inv
                const std::vector<int>& o = (*ord)[k];
                if ((int)o.size() != n[k]) return -1;
                ord_off_[k] = (int)aux.size();
                aux.insert(aux.end(), o.begin(), o.end());
                std::vector<int> inv(n[k], -1);
                for (int i = 0; i < n[k]; ++i) {
                    if (o[i] < 0 || o[i] >= n[k] || inv[o[i]] >= 0) return -1;
                    inv[o[i]] = i;
                }
                aux.insert(aux.end(), inv.begin(), inv.end());
            }
            if (env && !(*env)[k].last.empty()) {  // last, top, first per block
                const BlockEnvelope& e = (*env)[k];
                if ((int)e.last.size() != npad / kFB) return -1;
                env_off_[k] = (int)aux.size();
                aux.insert(aux.end(), e.last.begin(), e.last.end());
                aux.insert(aux.end(), e.top.begin(), e.top.end());
                aux.insert(aux.end(), e.first.begin(), e.first.end());
            }
            m.ord_off = ord_off_[k];
            m.env_off = env_off_[k];
            sd[k] = SolveDesc{a, n[k], npad, (int)b_off[k], (int)poff, ord_off_[k], env_off_[k]};
            a += (long long)npad * npad;
            poff += npad;
            btot = std::max(btot, b_off[k] + n[k]);
            npad_max_ = std::max(npad_max_, npad);
        }
        if (poff > INT_MAX || btot > INT_MAX || aux.size() > (size_t)INT_MAX) return -1;
        // the pivoted factorization composes each permutation in its LDS tile buffer
        if (npad_max_ > kPivTiles * kFB * kFLd * (int)(sizeof(double) / sizeof(int))) return -1;
        store_ = a;
        btot_ = btot;
        check_split_ = column_split((long long)(npad_max_ / kFB) * std::min(K, kGridY));
        // the factorization plan: matrices largest first, each within its envelope
        std::vector<int> order(K);
        for (int k = 0; k < K; ++k) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return md[x].npad > md[y].npad; });
        std::vector<FacDesc> fd;
        std::vector<const int*> last;
        for (int k : order) {
            fd.push_back({md[k].a_off, md[k].npad, 0});
            last.push_back(env_off_[k] >= 0 ? (*env)[k].last.data() : nullptr);
        }
        std::vector<int> pre;
        if (!plan_lu_steps(fd, last, steps_, pre)) return -1;
        lds_order_ = std::max(0, (lds / 8 - kSolveLds) / 256 * 256);
        const size_t scratch = npad_max_ <= lds_order_ ? 1 : (size_t)btot;
        if (hipMalloc(&fd_d_, sizeof(FacDesc) * fd.size()) != hipSuccess ||
            hipMalloc(&pre_d_, sizeof(int) * std::max<size_t>(pre.size(), 1)) != hipSuccess ||
            hipMalloc(&md_d_, sizeof(MatDesc) * K) != hipSuccess || hipMalloc(&sd_d_, sizeof(SolveDesc) * K) != hipSuccess ||
            hipMalloc(&perm_d_, sizeof(int) * (size_t)poff) != hipSuccess ||
            hipMalloc(&scratch_d_, sizeof(double) * scratch) != hipSuccess ||
            hipMalloc(&aux_d_, sizeof(int) * std::max<size_t>(aux.size(), 1)) != hipSuccess)
            return -2;
        if (hipMemcpy(fd_d_, fd.data(), sizeof(FacDesc) * fd.size(), hipMemcpyHostToDevice) != hipSuccess ||
            (!pre.empty() && hipMemcpy(pre_d_, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(md_d_, md.data(), sizeof(MatDesc) * K, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(sd_d_, sd.data(), sizeof(SolveDesc) * K, hipMemcpyHostToDevice) != hipSuccess ||
            (!aux.empty() && hipMemcpy(aux_d_, aux.data(), sizeof(int) * aux.size(), hipMemcpyHostToDevice) != hipSuccess))
            return -2;
        return 0;
    }

    long long a_off(int k) const override { return a_off_[k]; }
    long long store_size() const override { return store_; }
    const int* aux() const override { return aux_d_; }
    int ord_off(int k) const override { return ord_off_[k]; }
    int env_off(int k) const override { return env_off_[k]; }

    int factor_unpivoted(double* store, int* flag, hipStream_t st) override
    {
        flag_ = flag;  // the solves sweep the matrices refactored with pivoting whole
        if (d4::launch_fill_int(flag, 0, K_, st) != hipSuccess) return -2;
        if (launch_lu_steps(steps_, fd_d_, pre_d_, store, st)) return -2;
        for (int k0 = 0; k0 < K_; k0 += kGridY)
            hipLaunchKernelGGL(multiplier_check_kernel, dim3(npad_max_ / kFB * check_split_, std::min(kGridY, K_ - k0)),
                               dim3(256), 0, st, md_d_ + k0, store, flag + k0, aux_d_, check_split_);
        return hipGetLastError() != hipSuccess ? -2 : 0;
    }

    int factor_pivoted(double* store, const int* flag, int* singular, hipStream_t st) override
    {
        hipLaunchKernelGGL(lu_pivoted_kernel, dim3(K_), dim3(kPivThreads), 0, st, md_d_,
                           sd_d_, store, flag, perm_d_, singular);
        return hipGetLastError() != hipSuccess ? -2 : 0;
    }

    int solve(const double* store, double* b, hipStream_t st) override
    {
        if (npad_max_ <= lds_order_)
            hipLaunchKernelGGL(lu_solve_kernel<true>, dim3(K_), dim3(kSolveThreads),
                               sizeof(double) * (kSolveLds + npad_max_), st, sd_d_, store, perm_d_, b, scratch_d_,
                               aux_d_, flag_);
        else
            hipLaunchKernelGGL(lu_solve_kernel<false>, dim3(K_), dim3(kSolveThreads), sizeof(double) * kSolveLds,
                               st, sd_d_, store, perm_d_, b, scratch_d_, aux_d_, flag_);
        return hipGetLastError() != hipSuccess ? -2 : 0;
    }

private:
    static constexpr int kGridY = 65535;
    int K_ = 0, npad_max_ = 0, lds_order_ = 0;
    int check_split_ = 1;  // column groups per block of the multiplier check
    long long store_ = 0, btot_ = 0;
    std::vector<long long> a_off_;
    std::vector<int> ord_off_, env_off_;
    std::vector<LuStep> steps_;
    FacDesc* fd_d_ = nullptr;
    int* pre_d_ = nullptr;
    MatDesc* md_d_ = nullptr;
    SolveDesc* sd_d_ = nullptr;
    int* perm_d_ = nullptr;
    double* scratch_d_ = nullptr;
    int* aux_d_ = nullptr;
    const int* flag_ = nullptr;
};

}  // namespace

DenseLUGraph* new_dense_lu_graph(int device, int K, const int* n, const long long* b_off,
                                 const std::vector<std::vector<int>>* ord, const std::vector<BlockEnvelope>* env)
{
    GraphLU* g = new GraphLU();
    if (g->init(device, K, n, b_off, ord, env) != 0) {
        delete g;
        return nullptr;
    }
    return g;
}

LinearSolver* new_gpu_dense_lu(int device) { return new GpuDenseLU(device); }
BatchDenseLU* new_batch_dense_lu(int device) { return new BatchLU(device); }

}  // namespace d4host

void batch_lu_free(d4host::BatchDenseLU* lu) { delete lu; }

extern "C" int d4_lu_factor_dense_dev(d4_ctx* ctx, int n, const double* a, long long row_stride,
                                      long long col_stride, int* singular, int* unpivoted)
{
    if (!ctx) return D4_ERR_INVALID;
    if (n <= 0 || !a) return ctx->fail(D4_ERR_INVALID, "d4_lu_factor_dense_dev: n <= 0 or null matrix");
    if (!ctx->lu) ctx->lu = d4host::new_batch_dense_lu(ctx->device);
    if (hipDeviceSynchronize() != hipSuccess)  // the matrix may come from any stream
        return ctx->fail(D4_ERR_HIP, "d4_lu_factor_dense_dev: device synchronization failed");
    unsigned char s = 0, u = 0;
    const int rc = ctx->lu->factor_dense_dev(n, a, row_stride, col_stride, &s, &u);
    if (rc == -1) return ctx->fail(D4_ERR_INVALID, "d4_lu_factor_dense_dev: invalid order");
    if (rc < 0) return ctx->fail(D4_ERR_HIP, "d4_lu_factor_dense_dev: factorization failed on the GPU");
    if (singular) *singular = s;
    if (unpivoted) *unpivoted = u;
    return D4_OK;
}

extern "C" int d4_lu_solve_dense_dev(d4_ctx* ctx, int n, const double* b, double* x)
{
    if (!ctx) return D4_ERR_INVALID;
    if (!ctx->lu || !b || !x) return ctx->fail(D4_ERR_INVALID, "d4_lu_solve_dense_dev: nothing factored or null vector");
    if (hipDeviceSynchronize() != hipSuccess)
        return ctx->fail(D4_ERR_HIP, "d4_lu_solve_dense_dev: device synchronization failed");
    const int rc = ctx->lu->solve_dense_dev(n, b, x);
    if (rc == -1)
        return ctx->fail(D4_ERR_INVALID, "d4_lu_solve_dense_dev: n differs from the last d4_lu_factor_dense_dev");
    if (rc < 0) return ctx->fail(D4_ERR_HIP, "d4_lu_solve_dense_dev: solve failed on the GPU");
    return D4_OK;
}

extern "C" int d4_lu_solve_batch(d4_ctx* ctx, int K, const int* n, const int* const* Ap, const int* const* Ai,
                                 const double* const* Ax, const double* const* b, double* const* x,
                                 unsigned char* singular, unsigned char* unpivoted)
{
    if (!ctx) return D4_ERR_INVALID;
    if (K < 0 || (K > 0 && (!n || !Ap || !Ai || !Ax || !b || !x)))
        return ctx->fail(D4_ERR_INVALID, "d4_lu_solve_batch: null argument or K < 0");
    if (K == 0) return D4_OK;
    if (!ctx->lu) ctx->lu = d4host::new_batch_dense_lu(ctx->device);
    std::vector<d4host::BatchDenseLU::Matrix> m(K);
    for (int k = 0; k < K; ++k) m[k] = {n[k], Ap[k], Ai[k], Ax[k]};
    int rc = ctx->lu->factor(K, m.data(), singular, unpivoted);
    if (rc == -1) return ctx->fail(D4_ERR_INVALID, "d4_lu_solve_batch: invalid matrix (order or CSC structure)");
    if (rc < 0) return ctx->fail(D4_ERR_HIP, "d4_lu_solve_batch: factorization failed on the GPU");
    rc = ctx->lu->solve(K, b, x);
    if (rc < 0) return ctx->fail(rc == -1 ? D4_ERR_INVALID : D4_ERR_HIP, "d4_lu_solve_batch: solve failed");
    return D4_OK;
}
