// Supernodal triangular solves L z = r and L^T x = z on the factor panels (replace the
// cusparseSpSV analysis + solve pairs of src/cuda/kkt.cu:731-734,788-789, which re-ran the
// analysis on every call).  Levels of the supernodal etree are processed leaves-first (forward)
// and root-first (backward), one launch per level and one workgroup per supernode; inside a
// supernode, 64-column blocks are processed in order by the workgroup's wavefronts (see below).
//  * diagonal blocks: one wavefront, lane i keeps row (forward) or column (backward) i of the
//    row-major diagonal copy in registers; broadcasts by lane shuffles, no barriers;
//  * off-diagonal parts: coalesced GEMV (forward) or LDS-transposed GEMV^T (backward); rows
//    owned by ancestors are accumulated in a scratch vector and added with atomics (forward).
// Vectors are FP64 for both FP64 and FP32 panels.
#include "common.hpp"

namespace {

constexpr int TS = 64;
constexpr int DP = 65;

// Wave-uniform broadcast of lane `src`'s value (src is a compile-time constant at every use):
// two readlanes into scalar registers, no LDS round trip.
__device__ __forceinline__ double bcast_lane(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

// Stage the row-major copy Lr (kb x kb, reciprocal diagonal) of a diagonal block into Ds (64
// rows of pitch 65 doubles; zero past kb) in four batches of 16 rows: coalesced loads, issued
// ahead of the solve that reads them.
template <typename T>
__device__ __forceinline__ void stage_diag(const T* __restrict__ Lr, int kb, double* Ds) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        T v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = 16 * q + i;
            v[i] = (r < kb && lane < kb) ? Lr[r * kb + lane] : T(0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) Ds[(16 * q + i) * 65 + lane] = (double)v[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// L z = b on a staged diagonal block, one wavefront: lane i holds b_i and reads row i of L
// from LDS; z_j is formed in lane j (b_j times the reciprocal diagonal) and broadcast by
// readlane, so each step's dependency chain is one multiply, two readlanes and one fma.
__device__ __forceinline__ double fsub_staged(double b, const double* Ds, int kb) {
    const int lane = threadIdx.x & 63;
    const double* row = Ds + lane * 65;
    const double rd = lane < kb ? row[lane] : 0.0;
#pragma unroll
    for (int j = 0; j < TS; ++j) {
        if (j < kb) {
            const double zj = bcast_lane(b * rd, j);
            if (lane == j) b = zj;
            else if (lane > j) b -= row[j] * zj;
        }
    }
    return b;
}

// L^T x = r on a diagonal block held in registers: lane j holds r_j and col[i] = L[i][j] (row i
// of the row-major copy, element j: coalesced loads), rd = 1 / L[j][j]; x_i is formed in lane i
// and broadcast by readlane.
template <typename T>
__device__ __forceinline__ double bsub_regs(double r, const T (&col)[TS], double rd, int kb) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = TS - 1; i >= 0; --i) {
        if (i < kb) {
            const double xi = bcast_lane(r * rd, i);
            if (lane == i) r = xi;
            else if (lane < i) r -= (double)col[i] * xi;
        }
    }
    return r;
}

// ------------------------------------------------------------------------------------------
// Per-level solves: one workgroup of NW wavefronts per supernode of the level, one launch per
// level.  The row tiles of the panel (t < nb: the rows of diagonal block t; t >= nb: 64-row
// groups of the rows below the supernode, owned by ancestors) belong to wavefront t % NW.
// Wavefronts hand results to each other through workgroup-scope release / acquire on LDS
// progress words, so nobody waits at a full barrier and no device-scope fence is needed
// (device-scope acquire invalidates a whole L2 on this chip).
//  forward   for j = 0 .. nb-1: the owner of tile j solves its diagonal block (it applied all
//            updates of tile j itself), publishes z_j in an LDS ring of ZR blocks and advances
//            `done`; every wavefront applies L(t, j) z_j to its tiles t > j (diagonal tiles in
//            x, rows below in a scratch vector that is added to the ancestors' entries with
//            atomics at the end).  A ring slot is reused only after every wavefront finished
//            the iteration that read it (`prog`).
//  backward  for j = nb-1 .. 0: every wavefront forms part_v = sum over its tiles t > j of
//            L(t, j)^T x_t (LDS-transposed CW-column chunks, fixed reduction order), stores it
//            in `pbuf` and publishes; the owner of j waits for all NW parts, sums them in
//            wavefront order (deterministic), solves its diagonal block and keeps x_j.
// Task fields: aux = diagonal-copy offset of block 0 (block j at aux + 4096 j), aux2 = row-list
// offset, kbeg = offset of the rows below in `below`, ws = offset of the supernode's partial
// sums in `pbuf` (nb * NW * 64 doubles).  Waits are bounded; a timeout sets err bit 0x10.
constexpr long long kSpinLimit = 1ll << 26;
constexpr int ZR = 32;

struct Tiles {   // row tiles of a supernode panel (rows: 0 .. ld; columns: 0 .. nc)
    int64_t ld;
    int nc, nb, nt;
    __device__ Tiles(int64_t ld_, int nc_)
        : ld(ld_), nc(nc_), nb((nc_ + TS - 1) / TS), nt(nb + (int)((ld_ - nc_ + TS - 1) / TS)) {}
    __device__ __forceinline__ int64_t beg(int t) const { return t < nb ? (int64_t)t * TS : nc + (int64_t)(t - nb) * TS; }
    __device__ __forceinline__ int64_t end(int t) const {
        return t < nb ? min((int64_t)(t + 1) * TS, (int64_t)nc) : min(nc + (int64_t)(t - nb + 1) * TS, ld);
    }
};

// First tile t >= t0 owned by wavefront v.
template <int NW>
__device__ __forceinline__ int first_tile(int t0, int v) { return t0 + ((v - t0 % NW + NW) % NW); }

// Called by every lane of a wavefront after its data stores: release at workgroup scope, then
// lane 0 writes the progress word.
__device__ __forceinline__ void lds_publish(int* w, int val) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(w, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// Wave-uniform wait until pred(*w) holds; false on timeout (err bit 0x10 set).
template <typename Pred>
__device__ __forceinline__ bool lds_wait(const int* w, Pred pred, int32_t* err) {
    long long n = 0;
    while (!pred(__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))) {
        if (++n >= kSpinLimit) {
            if ((threadIdx.x & 63) == 0) atomicOr(err, 0x10);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}

// One 64-row tile x 64 columns of a panel, a column per register (lane = row): all loads are
// issued together (coalesced 512-byte rows of 64 lanes).  Columns >= kb and rows outside the
// tile read as zero.
template <typename T>
__device__ __forceinline__ void load_tile(T (&lv)[TS], const T* __restrict__ col, int64_t ld, int kb, bool row_ok) {
#pragma unroll
    for (int c = 0; c < TS; ++c) lv[c] = (row_ok && c < kb) ? col[(int64_t)c * ld] : T(0);
}

// x_t -= L(t, j) z_j for the 64 columns of block j held in lv (tile t: diagonal tiles update
// the supernode's own entries xs, tiles below update the scratch vector bl).
template <typename T>
__device__ __forceinline__ void fwd_apply(const Tiles tl, int t, const T (&lv)[TS], const double* z,
                                          double* xs, double* bl) {
    const int64_t row = tl.beg(t) + (threadIdx.x & 63);
    if (row >= tl.end(t)) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
    for (int c = 0; c < TS; c += 4) {
        a0 += (double)lv[c] * z[c];
        a1 += (double)lv[c + 1] * z[c + 1];
        a2 += (double)lv[c + 2] * z[c + 2];
        a3 += (double)lv[c + 3] * z[c + 3];
    }
    const double a = (a0 + a1) + (a2 + a3);
    if (t < tl.nb) xs[row] -= a;
    else bl[row - tl.nc] -= a;
}

// Solve diagonal block jb of the supernode (tile jb fully updated by the calling wavefront,
// its factor staged in Ds by the same wavefront), put z in x and in the LDS ring, publish
// `done`.  Ring slot reuse waits until every wavefront finished iteration jb - ZR.
template <int NW>
__device__ __forceinline__ bool fwd_solve_block(const lkd_task& tk, int jb, double* x, double* zs,
                                                const double* Ds, int* done, const int* prog,
                                                int32_t* err) {
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (jb >= ZR)
        for (int u = 0; u < NW; ++u)
            if (u != v && !lds_wait(&prog[u], [&](int p) { return p > jb - ZR; }, err)) return false;
    const int kb = min(TS, tk.ncol - jb * TS);
    double* xj = x + tk.first + jb * TS;
    const double b = lane < kb ? xj[lane] : 0.0;
    const double zj = fsub_staged(b, Ds, kb);
    if (lane < kb) xj[lane] = zj;
    zs[(jb % ZR) * TS + lane] = lane < kb ? zj : 0.0;
    lds_publish(done, jb + 1);
    return true;
}

template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) void k_fwd_sn(const lkd_task* __restrict__ tasks,
                                                    const T* __restrict__ L,
                                                    const T* __restrict__ diag,
                                                    const int32_t* __restrict__ sn_rows,
                                                    double* __restrict__ x,
                                                    double* __restrict__ below, int32_t* err) {
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ double zs[ZR * TS];                    // ring of solved blocks (zero past kb)
    __shared__ double Ds[2][TS * 65];                 // staged diagonal blocks (block jb: jb & 1)
    __shared__ int done;                              // blocks solved
    __shared__ int prog[NW];                          // iterations finished per wavefront
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const Tiles tl(ld, tk.ncol);
    const int nb = tl.nb, nt = tl.nt, nc = tk.ncol;
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double* bl = below + tk.kbeg;
    if (threadIdx.x == 0) done = 0;
    if (lane == 0) prog[v] = 0;
    for (int t = first_tile<NW>(nb, v); t < nt; t += NW)
        if (tl.beg(t) + lane < tl.end(t)) bl[tl.beg(t) + lane - nc] = 0.0;
    __syncthreads();
    auto stage = [&](int jb) {
        stage_diag<T>(diag + tk.aux + (int64_t)jb * TS * TS, min(TS, nc - jb * TS), Ds[jb & 1]);
    };
    if (v == 0) {
        stage(0);
        if (!fwd_solve_block<NW>(tk, 0, x, zs, Ds[0], &done, prog, err)) return;
    }
    for (int j = 0; j < nb; ++j) {
        const int kbj = min(TS, nc - j * TS);
        const T* Pj = P + (int64_t)j * TS * ld;
        const int t1 = first_tile<NW>(j + 1, v);
        // the owner of block j + 1 stages its factor now: slot (j + 1) & 1 was last read by the
        // solve of block j - 1, which finished before this wavefront passed iteration j - 1
        if (t1 == j + 1 && t1 < nb) stage(t1);
        T lv[TS];
        // the first tile's panel data does not depend on z_j: load it before waiting
        load_tile<T>(lv, Pj + tl.beg(min(t1, nt - 1)) + lane, ld, kbj, t1 < nt && tl.beg(t1) + lane < tl.end(t1));
        if (!lds_wait(&done, [&](int d) { return d > j; }, err)) return;
        const double* z = zs + (j % ZR) * TS;
        if (t1 < nt) {
            fwd_apply<T>(tl, t1, lv, z, x + tk.first, bl);
            if (t1 == j + 1 && t1 < nb && !fwd_solve_block<NW>(tk, t1, x, zs, Ds[t1 & 1], &done, prog, err))
                return;   // look-ahead: the next block first
        }
        for (int t = t1 + NW; t < nt; t += NW) {
            load_tile<T>(lv, Pj + tl.beg(t) + lane, ld, kbj, tl.beg(t) + lane < tl.end(t));
            fwd_apply<T>(tl, t, lv, z, x + tk.first, bl);
        }
        if (nb > ZR) lds_publish(&prog[v], j + 1);
    }
    for (int t = first_tile<NW>(nb, v); t < nt; t += NW) {
        const int64_t row = tl.beg(t) + lane;
        if (row < tl.end(t)) atomic_add(x + sn_rows[tk.aux2 + row], bl[row - nc]);
    }
}

template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) void k_bwd_sn(const lkd_task* __restrict__ tasks,
                                                    const T* __restrict__ L,
                                                    const T* __restrict__ diag,
                                                    const int32_t* __restrict__ sn_rows,
                                                    double* __restrict__ x,
                                                    double* __restrict__ below,
                                                    double* __restrict__ pbuf, int32_t* err) {
    constexpr int CW = 16;                            // columns per transposed chunk
    constexpr int NQ = TS / CW;
    const lkd_task tk = tasks[blockIdx.x];
    __shared__ double Ts[NW][CW * 65];                // per wavefront: CW columns x 64 rows
    __shared__ int pub[NW];                           // lowest block each wavefront published
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const Tiles tl(ld, tk.ncol);
    const int nb = tl.nb, nt = tl.nt, nc = tk.ncol;
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cl = lane % CW, rg = lane / CW;         // column of a chunk, row group
    double* bl = below + tk.kbeg;
    double* pb = pbuf + tk.ws;                        // [nb][NW][64]
    double* Tw = Ts[v];
    for (int t = first_tile<NW>(nb, v); t < nt; t += NW) {
        const int64_t row = tl.beg(t) + lane;
        if (row < tl.end(t)) bl[row - nc] = x[sn_rows[tk.aux2 + row]];
    }
    if (lane == 0) pub[v] = nb;
    __syncthreads();
    // part[q] (lanes cl): += sum over the rows of tile t of L(t, jb)^T x_t, columns CW q + cl.
    auto contrib = [&, tl](double (&part)[NQ], int t, int jb) {
        const int kb = min(TS, nc - jb * TS);
        const int64_t r0 = tl.beg(t);
        const bool ok = r0 + lane < tl.end(t);
        const double xr = ok ? (t < nb ? x[tk.first + r0 + lane] : bl[r0 + lane - nc]) : 0.0;
        T lv[TS];
        load_tile<T>(lv, P + (int64_t)jb * TS * ld + r0 + lane, ld, kb, ok);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
#pragma unroll
            for (int c = 0; c < CW; ++c) Tw[c * 65 + lane] = (double)lv[CW * q + c] * xr;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < CW; ++i) s += Tw[cl * 65 + CW * rg + i];
#pragma unroll
            for (int o = CW; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
            part[q] += s;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    };
    double pre[NQ];          // look-ahead contributions to block pre_blk (tiles t > pre_blk + 1)
    int pre_blk = -1;
    for (int j = nb - 1; j >= 0; --j) {
        double part[NQ];
        const bool have = pre_blk == j;
#pragma unroll
        for (int q = 0; q < NQ; ++q) part[q] = have ? pre[q] : 0.0;
        if (have) contrib(part, j + 1, j);            // the tile solved last is the only one left
        else
            for (int t = first_tile<NW>(j + 1, v); t < nt; t += NW) contrib(part, t, j);
        double* pj = pb + ((int64_t)j * NW + v) * TS;
        if (lane < CW) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) pj[CW * q + lane] = part[q];
        }
        lds_publish(&pub[v], j);
        if (j % NW != v) continue;
        // owner of block j: while the other parts arrive, start block j - 1 with the tiles
        // that do not need x_j
        if (j > 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) pre[q] = 0.0;
            for (int t = first_tile<NW>(j + 1, v); t < nt; t += NW) contrib(pre, t, j - 1);
            pre_blk = j - 1;
        }
        // the diagonal block's columns load while the other parts arrive
        const int kbj = min(TS, nc - j * TS);
        const T* Lr = diag + tk.aux + (int64_t)j * TS * TS;
        T col[TS];
#pragma unroll
        for (int i = 0; i < TS; ++i) col[i] = (i < kbj && lane < kbj) ? Lr[i * kbj + lane] : T(0);
        const double rd = lane < kbj ? (double)Lr[lane * kbj + lane] : 0.0;
        for (int u = 0; u < NW; ++u)
            if (u != v && !lds_wait(&pub[u], [&](int p) { return p <= j; }, err)) return;
        double val = 0.0;
        if (lane < kbj) {
            double s = 0.0;
            const double* p0 = pb + (int64_t)j * NW * TS + lane;
            for (int u = 0; u < NW; ++u) s += p0[(int64_t)u * TS];
            val = x[tk.first + j * TS + lane] - s;
        }
        const double sol = bsub_regs<T>(val, col, rd, kbj);
        if (lane < kbj) x[tk.first + j * TS + lane] = sol;
    }
}

template <typename T, int NW>
int level_solve(int dir, int32_t ntask, const lkd_task* tasks, const void* panel, const void* diag,
                const int32_t* rows, double* x, double* below, double* pbuf, int32_t* err,
                void* stream) {
    if (dir == 0)
        LK_LAUNCH((k_fwd_sn<T, NW>), ntask, 64 * NW, 0, stream, tasks, (const T*)panel,
                  (const T*)diag, rows, x, below, err);
    else
        LK_LAUNCH((k_bwd_sn<T, NW>), ntask, 64 * NW, 0, stream, tasks, (const T*)panel,
                  (const T*)diag, rows, x, below, pbuf, err);
    return lk_last_error();
}

template <typename T>
int level_solve_nw(int nw, int dir, int32_t ntask, const lkd_task* tasks, const void* panel,
                   const void* diag, const int32_t* rows, double* x, double* below, double* pbuf,
                   int32_t* err, void* stream) {
    return nw == 8 ? level_solve<T, 8>(dir, ntask, tasks, panel, diag, rows, x, below, pbuf, err, stream)
                    : level_solve<T, 4>(dir, ntask, tasks, panel, diag, rows, x, below, pbuf, err, stream);
}

}  // namespace

extern "C" {

int lkd_trsv_level(int prec, int dir, int nw, int32_t ntask, const lkd_task* tasks,
                   const void* panel, const void* diag, const int32_t* sn_rows, double* x,
                   double* below, double* pbuf, int32_t* err, void* stream) {
    if (ntask <= 0) return 0;
    if (nw != 4 && nw != 8) return (int)hipErrorInvalidValue;
    return prec == LKD_F32
               ? level_solve_nw<float>(nw, dir, ntask, tasks, panel, diag, sn_rows, x, below, pbuf, err, stream)
               : level_solve_nw<double>(nw, dir, ntask, tasks, panel, diag, sn_rows, x, below, pbuf, err, stream);
}

}  // extern "C"
