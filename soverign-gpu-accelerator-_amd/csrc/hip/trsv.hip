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
// ahead of the solve that reads them.  Scaled: every entry below the diagonal is then divided
// by the diagonal of its column (lane = column), Ds[r][c] = L[r][c] / L[c][c] (fsub_scaled).
template <typename T, bool Scaled = false>
__device__ __forceinline__ void stage_diag(const T* __restrict__ Lr, int kb, double* Ds) {
    const int lane = threadIdx.x & 63;
    if (kb == TS) {                               // full block: unconditional loads
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            T v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = Lr[(16 * q + i) * TS + lane];
#pragma unroll
            for (int i = 0; i < 16; ++i) Ds[(16 * q + i) * 65 + lane] = (double)v[i];
        }
    } else {                                      // clamped loads times a 0/1 mask: no branches
        const int lc = min(lane, kb - 1);
        for (int r = 0; r < TS; ++r)
            Ds[r * 65 + lane] = (double)Lr[min(r, kb - 1) * kb + lc] * ((r < kb && lane < kb) ? 1.0 : 0.0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!Scaled) return;
    const double rc = Ds[lane * 65 + lane];
#pragma unroll 8
    for (int r = 0; r < TS; ++r) Ds[r * 65 + lane] *= r > lane ? rc : 1.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// L z = b on a staged diagonal block, one wavefront.  Lane i keeps b_i and its row of L in
// registers with the diagonal entry zeroed (the row-major copy is zero above the diagonal), so
// step j is b -= L[i][j] z_j for every lane at once: lanes <= j are left unchanged by the zero
// entries, lane j keeps its updated right-hand side and z = b / L_ii is formed at the end.
// z_j itself comes from lane j by readlane: the dependency chain per step is one multiply, two
// readlanes and one fma, with no branch and no memory access.  Past kb the staged rows and the
// reciprocal diagonal are zero, so all 64 steps run for any kb.
__device__ __forceinline__ double fsub_staged(double b, const double* Ds, int kb) {
    const int lane = threadIdx.x & 63;
    const double* row = Ds + lane * 65;
    double rv[TS];
#pragma unroll
    for (int j = 0; j < TS; ++j) rv[j] = row[j];
    const double rd = lane < kb ? row[lane] : 0.0;
#pragma unroll
    for (int j = 0; j < TS; ++j) rv[j] = lane == j ? 0.0 : rv[j];
#pragma unroll
    for (int j = 0; j < TS; ++j) b = fma(-rv[j], bcast_lane(b * rd, j), b);
    return b * rd;
}

// fsub_staged on a block staged with Scaled: the entries are M[i][j] = L[i][j] / L[j][j], so
// step j is b -= M[i][j] b_j (lane j keeps b_j = L_jj z_j) and the multiply leaves the chain:
// two readlanes and one fma per step.
__device__ __forceinline__ double fsub_scaled(double b, const double* Ds, int kb) {
    const int lane = threadIdx.x & 63;
    const double* row = Ds + lane * 65;
    double rv[TS];
#pragma unroll
    for (int j = 0; j < TS; ++j) rv[j] = row[j];
    const double rd = lane < kb ? row[lane] : 0.0;
#pragma unroll
    for (int j = 0; j < TS; ++j) rv[j] = lane == j ? 0.0 : rv[j];
#pragma unroll
    for (int j = 0; j < TS; ++j) b = fma(-rv[j], bcast_lane(b, j), b);
    return b * rd;
}

// L^T x = r on a diagonal block held in registers: lane j holds r_j and col[i] = L[i][j] (row i
// of the row-major copy, element j: coalesced loads), rd = 1 / L[j][j]; x_i is formed in lane i
// and broadcast by readlane.
// Same scheme as fsub_staged: col[lane] (the diagonal) must be zero and col[i] = 0 for i < lane
// (upper part of the copy), so step i is r -= L[i][j] x_i for every lane; x = r / L_jj at the end.
template <typename T>
__device__ __forceinline__ double bsub_regs(double r, const T (&col)[TS], double rd) {
#pragma unroll
    for (int i = TS - 1; i >= 0; --i) r = fma(-(double)col[i], bcast_lane(r * rd, i), r);
    return r * rd;
}

// ------------------------------------------------------------------------------------------
// Per-level solves: one launch per level, G workgroups of NW wavefronts per supernode (G = 1
// unless the level has very few supernodes).  The row tiles of the panel (t < nb: the rows of
// diagonal block t; t >= nb: 64-row groups of the rows below the supernode, owned by ancestors)
// belong to the wavefront with residue t % (G NW), where wavefront v of workgroup g has residue
// g + G v (see Own).  Inside a workgroup, wavefronts hand results to each other through
// workgroup-scope release / acquire on LDS progress words, so nobody waits at a full barrier;
// between the workgroups of a supernode, which share one XCD's L2, a flag in device memory is
// set after the data stores completed and the reader invalidates its L1 (dev_publish /
// dev_wait), once per hand-off.
//  forward   for j = 0 .. nb-1: the owner of tile j solves its diagonal block (it applied all
//            updates of tile j itself), publishes z_j in its workgroup's LDS ring of ZR blocks
//            (and, G > 1, in device memory, where wavefront 0 of every other workgroup picks it
//            up) and advances `done`; every wavefront applies L(t, j) z_j to its tiles t > j
//            (diagonal tiles in x, rows below in a scratch vector that is added to the
//            ancestors' entries with atomics at the end).  A ring slot is reused only after
//            every wavefront of the workgroup finished the iteration that read it (`prog`).
//  backward  (G = 1 or 2) for j = nb-1 .. 0: every wavefront forms part_v = sum over its
//            tiles t > j of L(t, j)^T x_t (LDS-transposed CW-column chunks, fixed reduction
//            order), stores it in `pbuf` (G = 1, the owner of j + 1: in LDS) and publishes; the
//            owner of j waits for all parts, sums them in residue order (deterministic), solves
//            its diagonal block and keeps x_j.  Four workgroups measured slower than two.
// Task fields: aux = diagonal-copy offset of block 0 (block j at aux + 4096 j), aux2 = row-list
// offset, kbeg = offset of the rows below in `below`, r0 = g, c0 = G, ws = offset of the
// supernode's area in `pbuf`, px = offset of its flags.  Waits are bounded; a timeout sets
// err bit 0x10.
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
// issued together (coalesced 512-byte rows of 64 lanes).  `col` must point at a row inside the
// panel for every lane (callers clamp it); columns >= kb and lanes with !row_ok read as zero.
// Full blocks (kb = 64, all but a supernode's last) take a path without any column condition.
template <typename T>
__device__ __forceinline__ void load_tile(T (&lv)[TS], const T* __restrict__ col, int64_t ld, int kb, bool row_ok) {
    if (kb == TS) {                               // full block: unconditional loads
#pragma unroll
        for (int c = 0; c < TS; ++c) {
            const T e = col[(int64_t)c * ld];
            lv[c] = row_ok ? e : T(0);
        }
    } else {                                      // clamped loads times a 0/1 mask: no branches
#pragma unroll
        for (int c = 0; c < TS; ++c)
            lv[c] = col[(int64_t)min(c, kb - 1) * ld] * ((row_ok && c < kb) ? T(1) : T(0));
    }
}

// Half a tile (32 columns c0 .. c0+31) into registers, same conventions as load_tile.
template <typename T>
__device__ __forceinline__ void load_half(T (&h)[TS / 2], const T* __restrict__ col, int64_t ld, int c0,
                                          int kb, bool row_ok) {
    if (kb == TS) {
#pragma unroll
        for (int c = 0; c < TS / 2; ++c) {
            const T e = col[(int64_t)(c0 + c) * ld];
            h[c] = row_ok ? e : T(0);
        }
    } else {                                      // clamped loads times a 0/1 mask: no branches
#pragma unroll
        for (int c = 0; c < TS / 2; ++c) {
            const int cc = c0 + c;
            h[c] = col[(int64_t)min(cc, kb - 1) * ld] * ((row_ok && cc < kb) ? T(1) : T(0));
        }
    }
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

// Tile ownership among the G workgroups of a supernode (G * NW wavefronts; wavefront v of
// workgroup g has the residue gw = g + G v): tile t belongs to the wavefront with residue
// t % (G * NW), so consecutive tiles go to consecutive workgroups and the rows of a supernode
// are spread evenly over them (every forward hand-off then crosses workgroups, in one L2).
template <int G, int NW>
struct Own {
    static constexpr int nwt = G * NW;
    int gw, g;
    __device__ __forceinline__ int first(int t0) const { return t0 + ((gw - t0 % nwt + nwt) % nwt); }
    __device__ __forceinline__ int wg_of(int t) const { return (t % nwt) % G; }
};

// Hand-off between the workgroups of one supernode, which all run on one XCD (they take their
// tasks from that XCD's queue, see xcd_task): the writing wavefront waits until its data
// stores have reached the shared L2, then sets the flag (relaxed); the reader spins on the flag
// and invalidates its compute unit's L1 before reading.  Flags hold the solve's epoch, so they
// never need resetting.
__device__ __forceinline__ void dev_publish(uint32_t* flag, uint32_t ep) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(flag, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool dev_wait(const uint32_t* flag, uint32_t ep, int32_t* err) {
    long long n = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ep) {
        if (++n >= kSpinLimit) {
            if ((threadIdx.x & 63) == 0) atomicOr(err, 0x10);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    asm volatile("buffer_inv sc0" ::: "memory");
    return true;
}

// Task of this workgroup in a launch with several workgroups per supernode: the next ticket of
// the queue of the XCD it runs on (XCC_ID mod nx), queue[qoff[x] .. qoff[x+1]); -1 when that
// queue is empty.  A supernode's workgroups are consecutive in one queue, and tickets go to
// running workgroups in order, so the earliest unfinished supernode has all its workgroups
// running (the XCD holds at least G of them at once).
struct XcdQueue {
    const int32_t* queue;
    const int32_t* qoff;
    uint32_t* cnt;
    int nx;
};
__device__ __forceinline__ int xcd_task(const XcdQueue& q, int* s_ti) {
    if (threadIdx.x == 0) {
        unsigned xcc;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
        const int x = (int)(xcc % (unsigned)q.nx);
        const int t = (int)atomicAdd(q.cnt + x, 1u);
        *s_ti = t < q.qoff[x + 1] - q.qoff[x] ? q.queue[q.qoff[x] + t] : -1;
    }
    __syncthreads();
    return *s_ti;
}

// Ring slot of block jb may be (re)written once every wavefront of the workgroup finished
// iteration jb - ZR.
template <int NW>
__device__ __forceinline__ bool ring_free(int jb, const int* prog, int32_t* err) {
    const int v = threadIdx.x >> 6;
    if (jb >= ZR)
        for (int u = 0; u < NW; ++u)
            if (u != v && !lds_wait(&prog[u], [&](int p) { return p > jb - ZR; }, err)) return false;
    return true;
}

// Solve diagonal block jb of the supernode (tile jb fully updated by the calling wavefront,
// its factor staged in Ds by the same wavefront), put z in x and in the LDS ring, publish
// `done`; with several workgroups also in the supernode's exchange area zg and its flag.
template <int NW, int G>
__device__ __forceinline__ bool fwd_solve_block(const lkd_task& tk, int jb, double* x, double* zs,
                                                const double* Ds, int* done, const int* prog,
                                                double* zg, uint32_t* fl, uint32_t ep,
                                                int32_t* err) {
    const int lane = threadIdx.x & 63;
    if (!ring_free<NW>(jb, prog, err)) return false;
    const int kb = min(TS, tk.ncol - jb * TS);
    double* xj = x + tk.first + jb * TS;
    const double b = lane < kb ? xj[lane] : 0.0;
    const double zj = fsub_staged(b, Ds, kb);
    const double zv = lane < kb ? zj : 0.0;
    if (lane < kb) xj[lane] = zj;
    zs[(jb % ZR) * TS + lane] = zv;
    if (G > 1) {
        zg[jb * TS + lane] = zv;
        dev_publish(fl + jb, ep);
    }
    lds_publish(done, jb + 1);
    return true;
}

// Forward level solve.  Task: r0 = g, c0 = G (workgroups of the supernode), ws = exchange area
// in zbuf (nb * 64 doubles), px = flag area (nb flags); see above for the other fields.
template <typename T, int NW, int G>
__global__ __launch_bounds__(64 * NW) void k_fwd_sn(const lkd_task* __restrict__ tasks,
                                                    const T* __restrict__ L,
                                                    const T* __restrict__ diag,
                                                    const int32_t* __restrict__ sn_rows,
                                                    double* __restrict__ x,
                                                    double* __restrict__ below, double* zbuf,
                                                    uint32_t* flags, const uint32_t* epoch,
                                                    int32_t* err, XcdQueue xq) {
    __shared__ int s_ti;
    const int ti = G > 1 ? xcd_task(xq, &s_ti) : (int)blockIdx.x;
    if (ti < 0) return;
    const lkd_task tk = tasks[ti];
    __shared__ double zs[ZR * TS];                    // ring of solved blocks (zero past kb)
    __shared__ double Ds[2][TS * 65];                 // staged diagonal blocks (block jb: jb & 1)
    __shared__ int done;                              // blocks available in the ring
    __shared__ int prog[NW];                          // iterations finished per wavefront
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const Tiles tl(ld, tk.ncol);
    const int nb = tl.nb, nt = tl.nt, nc = tk.ncol;
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const Own<G, NW> own{tk.r0 + G * v, tk.r0};
    const uint32_t ep = G > 1 ? *epoch : 0u;
    double* zg = zbuf + tk.ws;
    uint32_t* fl = flags + tk.px;
    double* bl = below + tk.kbeg;
    if (threadIdx.x == 0) done = 0;
    if (lane == 0) prog[v] = 0;
    for (int t = own.first(nb); t < nt; t += own.nwt)
        if (tl.beg(t) + lane < tl.end(t)) bl[tl.beg(t) + lane - nc] = 0.0;
    __syncthreads();
    auto stage = [&](int jb) {
        stage_diag<T>(diag + tk.aux + (int64_t)jb * TS * TS, min(TS, nc - jb * TS), Ds[jb & 1]);
    };
    if (own.gw == 0) {
        stage(0);
        if (!fwd_solve_block<NW, G>(tk, 0, x, zs, Ds[0], &done, prog, zg, fl, ep, err)) return;
    }
#pragma unroll 1
    for (int j = 0; j < nb; ++j) {
        const int kbj = min(TS, nc - j * TS);
        const T* Pj = P + (int64_t)j * TS * ld;
        const int t1 = own.first(j + 1);
        // the owner of block j + 1 stages its factor now: slot (j + 1) & 1 was last read by the
        // solve of an earlier block of this workgroup, finished before iteration j - 1 ended here
        if (t1 == j + 1 && t1 < nb) stage(t1);
        T lv[TS];
        // the first tile's panel data does not depend on z_j: load it before waiting
        {
            const int tt = min(t1, nt - 1);
            load_tile<T>(lv, Pj + min(tl.beg(tt) + lane, tl.end(tt) - 1), ld, kbj,
                         t1 < nt && tl.beg(t1) + lane < tl.end(t1));
        }
        // z_j solved in another workgroup: wavefront 0 brings it into this workgroup's ring
        if (G > 1 && v == 0 && own.wg_of(j) != own.g) {
            if (!ring_free<NW>(j, prog, err) || !dev_wait(fl + j, ep, err)) return;
            zs[(j % ZR) * TS + lane] = zg[j * TS + lane];
            lds_publish(&done, j + 1);
        }
        if (!lds_wait(&done, [&](int d) { return d > j; }, err)) return;
        const double* z = zs + (j % ZR) * TS;
        if (t1 < nt) {
            fwd_apply<T>(tl, t1, lv, z, x + tk.first, bl);
            if (t1 == j + 1 && t1 < nb &&
                !fwd_solve_block<NW, G>(tk, t1, x, zs, Ds[t1 & 1], &done, prog, zg, fl, ep, err))
                return;   // look-ahead: the next block first
        }
#pragma unroll 1
        for (int t = t1 + own.nwt; t < nt; t += own.nwt) {
            load_tile<T>(lv, Pj + min(tl.beg(t) + lane, tl.end(t) - 1), ld, kbj, tl.beg(t) + lane < tl.end(t));
            fwd_apply<T>(tl, t, lv, z, x + tk.first, bl);
        }
        if (nb > ZR) lds_publish(&prog[v], j + 1);
    }
    for (int t = own.first(nb); t < nt; t += own.nwt) {
        const int64_t row = tl.beg(t) + lane;
        if (row < tl.end(t)) atomic_add(x + sn_rows[tk.aux2 + row], bl[row - nc]);
    }
}

// Backward level solve, G = 1 or 2 workgroups per supernode (task r0 = g, c0 = G; with G = 2
// from per-XCD queues, like the forward).  ws = partial sums in pbuf (nb * G * NW * 64 doubles:
// [block][wavefront residue][64], see Own).  Block j's owner sums the parts of all wavefronts
// in residue order.  The part of the owner of block j + 1 is on the critical path (it needs
// x_{j+1}): that wavefront computes everything else early (look-ahead), keeps x_{j+1} in
// registers after its solve, has the first half of tile (j + 1, j) in registers already (FP32
// panels; FP64 ones have no registers to spare) and, with G = 1, hands the part over in LDS
// (`cp`, two slots: slot j & 1 is rewritten only after block j + 2's owner read it, which
// precedes the solve of j + 1).  The other parts go through pbuf; with G = 2 the wavefronts of
// the workgroup that does not own block j publish theirs with one flag per block and workgroup
// (each waits for its own stores first, the last one to count itself sets the flag).
template <typename T, int NW, int G>
__global__ __launch_bounds__(64 * NW) void k_bwd_sn(const lkd_task* __restrict__ tasks,
                                                    const T* __restrict__ L,
                                                    const T* __restrict__ diag,
                                                    const int32_t* __restrict__ sn_rows,
                                                    double* __restrict__ x,
                                                    double* __restrict__ below,
                                                    double* pbuf, uint32_t* flags,
                                                    const uint32_t* epoch, int32_t* err,
                                                    XcdQueue xq) {
    constexpr int CW = 16;                            // columns per transposed chunk
    constexpr int NQ = TS / CW;
    constexpr int NWT = G * NW;                       // wavefronts of the supernode
    constexpr int NC = 64;                            // part counters (blocks in flight < NC)
    __shared__ int s_ti;
    const int ti = G > 1 ? xcd_task(xq, &s_ti) : (int)blockIdx.x;
    if (ti < 0) return;
    const lkd_task tk = tasks[ti];
    __shared__ double Ts[NW][CW * 65];                // per wavefront: CW columns x 64 rows
    __shared__ double cp[2][TS];                      // critical-path parts, slot j & 1 (G = 1)
    __shared__ int pub[NW];                           // lowest block each wavefront published
    __shared__ int cnt[NC];                           // G > 1: parts stored per block (mod NC)
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const Tiles tl(ld, tk.ncol);
    const int nb = tl.nb, nt = tl.nt, nc = tk.ncol;
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const Own<G, NW> own{tk.r0 + G * v, tk.r0};
    const uint32_t ep = G > 1 ? *epoch : 0u;
    uint32_t* fl = flags + tk.px + nb;                // G > 1: [block][workgroup]
    const int cl = lane % CW, rg = lane / CW;         // column of a chunk, row group
    double* bl = below + tk.kbeg;
    double* pb = pbuf + tk.ws;                        // [nb][NWT][64], by wavefront residue
    double* Tw = Ts[v];
    if (G > 1)
        for (int i = threadIdx.x; i < NC; i += 64 * NW) cnt[i] = 0;
    for (int t = own.first(nb); t < nt; t += NWT) {
        const int64_t row = tl.beg(t) + lane;
        if (row < tl.end(t)) bl[row - nc] = x[sn_rows[tk.aux2 + row]];
    }
    if (lane == 0) pub[v] = nb;
    __syncthreads();
    // part[q] (lanes cl): += sum over the rows of tile t of L(t, jb)^T x_t, columns CW q + cl,
    // for the tiles t = t0, t0 + NW, ... < tlim; software-pipelined by halves (32 columns): the
    // next half's loads are in flight while the current one goes through the LDS transpose.
    auto chunk = [&](double& acc, const T* h, double xr) {
#pragma unroll
        for (int c = 0; c < CW; ++c) Tw[c * 65 + lane] = (double)h[c] * xr;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < CW; ++i) s += Tw[cl * 65 + CW * rg + i];
#pragma unroll
        for (int o = CW; o < 64; o <<= 1) s += __shfl_xor(s, o, 64);
        acc += s;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto rowp = [&](const T* Pb, int u) { return Pb + min(tl.beg(u) + lane, tl.end(u) - 1); };
    auto rok = [&](int u) { return tl.beg(u) + lane < tl.end(u); };
    auto contrib_list = [&, tl](double (&part)[NQ], int t0, int jb, int tlim) {
        if (t0 >= tlim) return;
        const int kb = min(TS, nc - jb * TS);
        const T* Pb = P + (int64_t)jb * TS * ld;
        auto xval = [&](int u) {
            const int64_t r = tl.beg(u) + lane;
            return r < tl.end(u) ? (u < nb ? x[tk.first + r] : bl[r - nc]) : 0.0;
        };
        T h0[TS / 2], h1[TS / 2];
        load_half<T>(h0, rowp(Pb, t0), ld, 0, kb, rok(t0));
        load_half<T>(h1, rowp(Pb, t0), ld, TS / 2, kb, rok(t0));
#pragma unroll 1
        for (int t = t0; t < tlim; t += NWT) {
            const int tn = t + NWT < tlim ? t + NWT : t;  // past the end: reload, unused
            const double xr = xval(t);
            chunk(part[0], h0, xr);
            chunk(part[1], h0 + CW, xr);
            load_half<T>(h0, rowp(Pb, tn), ld, 0, kb, rok(tn));
            chunk(part[2], h1, xr);
            chunk(part[3], h1 + CW, xr);
            load_half<T>(h1, rowp(Pb, tn), ld, TS / 2, kb, rok(tn));
        }
    };
    double pre[NQ];          // look-ahead part for block pre_blk (tiles t > pre_blk + 1)
    int pre_blk = -1;
    double xlast = 0.0;      // x of the block this wavefront solved last (lane = row; 0 past kb)
    constexpr bool kEarly = sizeof(T) == 4;
    T hc[TS / 2];            // first half of tile (j, j - 1), loaded before the solve of block j
    for (int j = nb - 1; j >= 0; --j) {
        if (pre_blk == j) {
            // critical path: tile j + 1 of block j with x_{j+1} from registers, into LDS
            const int kb = min(TS, nc - j * TS);
            const T* Pb = P + (int64_t)j * TS * ld;
            T h1[TS / 2];
            if (!kEarly) load_half<T>(hc, rowp(Pb, j + 1), ld, 0, kb, rok(j + 1));
            load_half<T>(h1, rowp(Pb, j + 1), ld, TS / 2, kb, rok(j + 1));
            chunk(pre[0], hc, xlast);
            chunk(pre[1], hc + CW, xlast);
            chunk(pre[2], h1, xlast);
            chunk(pre[3], h1 + CW, xlast);
            double* dst = G == 1 ? cp[j & 1] : pb + ((int64_t)j * NWT + own.gw) * TS;
            if (lane < CW) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) dst[CW * q + lane] = pre[q];
            }
        } else {
            double part[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) part[q] = 0.0;
            contrib_list(part, own.first(j + 1), j, nt);
            double* pj = pb + ((int64_t)j * NWT + own.gw) * TS;
            if (lane < CW) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) pj[CW * q + lane] = part[q];
            }
        }
        lds_publish(&pub[v], j);
        if (G > 1 && own.wg_of(j) != own.g) {
            // every wavefront's part has reached the L2 before it counts itself; the last one
            // of the workgroup publishes them to the owner's workgroup
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            int last = 0;
            if (lane == 0) last = (atomicAdd(&cnt[j % NC], 1) + 1) % NW == 0;
            if (__builtin_amdgcn_readfirstlane(last)) dev_publish(fl + (int64_t)j * G + own.g, ep);
        }
        if ((j % NWT) != own.gw) continue;
        // owner of block j: while the other parts arrive, start block j - 1 with the tiles
        // that do not need x_j, and load the first half of tile (j, j - 1)
        if (j > 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) pre[q] = 0.0;
            contrib_list(pre, own.first(j + 1), j - 1, nt);
            pre_blk = j - 1;
            if (kEarly)
                load_half<T>(hc, rowp(P + (int64_t)(j - 1) * TS * ld, j), ld, 0, min(TS, nc - (j - 1) * TS), rok(j));
        }
        __builtin_amdgcn_sched_barrier(0);
        // the diagonal block's columns load while the other parts arrive
        const int kbj = min(TS, nc - j * TS);
        const T* Lr = diag + tk.aux + (int64_t)j * TS * TS;
        T col[TS];
        if (kbj == TS) {
#pragma unroll
            for (int i = 0; i < TS; ++i) col[i] = Lr[i * TS + lane];
        } else {
#pragma unroll
            for (int i = 0; i < TS; ++i)       // clamped loads times a 0/1 mask: no branches
                col[i] = Lr[min(i, kbj - 1) * kbj + min(lane, kbj - 1)] * ((i < kbj && lane < kbj) ? T(1) : T(0));
        }
#pragma unroll
        for (int i = 0; i < TS; ++i) col[i] = i == lane ? T(0) : col[i];
        const int ld_ = min(lane, kbj - 1);
        const double rd = lane < kbj ? (double)Lr[ld_ * kbj + ld_] : 0.0;
        const double rhs = lane < kbj ? x[tk.first + j * TS + lane] : 0.0;
        // the parts in pbuf load while the critical-path part is still being formed; all
        // parts are then summed in wavefront order
        double val = 0.0;
        if constexpr (G == 1) {
            const int prod = j + 1 < nb ? (j + 1) % NW : -1;    // wavefront whose part is in cp
            for (int u = 0; u < NW; ++u)
                if (u != v && u != prod && !lds_wait(&pub[u], [&](int p) { return p <= j; }, err)) return;
            double pv[NW];
            const double* p0 = pb + (int64_t)j * NW * TS + lane;
#pragma unroll
            for (int u = 0; u < NW; ++u) pv[u] = u == prod ? 0.0 : p0[(int64_t)u * TS];
            if (prod >= 0 && !lds_wait(&pub[prod], [&](int p) { return p <= j; }, err)) return;
            if (lane < kbj) {
                double s = 0.0;
#pragma unroll
                for (int u = 0; u < NW; ++u) s += u == prod ? cp[j & 1][lane] : pv[u];
                val = rhs - s;
            }
        } else {
            for (int u = 0; u < NW; ++u)
                if (u != v && !lds_wait(&pub[u], [&](int p) { return p <= j; }, err)) return;
            for (int h = 0; h < G; ++h)
                if (h != own.g && !dev_wait(fl + (int64_t)j * G + h, ep, err)) return;
            if (lane < kbj) {
                double s = 0.0;
                const double* p0 = pb + (int64_t)j * NWT * TS + lane;
                for (int r = 0; r < NWT; ++r) s += p0[(int64_t)r * TS];
                val = rhs - s;
            }
        }
        const double sol = bsub_regs<T>(val, col, rd);
        if (lane < kbj) x[tk.first + j * TS + lane] = sol;
        xlast = lane < kbj ? sol : 0.0;
    }
}

template <typename T, int NW, int G>
int level_solve(int dir, int32_t ntask, const lkd_task* tasks, const void* panel, const void* diag,
                const int32_t* rows, double* x, double* below, double* pbuf, uint32_t* flags,
                const uint32_t* epoch, int32_t* err, const XcdQueue& xq, int grid, void* stream) {
    if (dir == 0) {
        if constexpr (G == 1 || G == 4)
            LK_LAUNCH((k_fwd_sn<T, NW, G>), G > 1 ? grid : ntask, 64 * NW, 0, stream, tasks, (const T*)panel,
                      (const T*)diag, rows, x, below, pbuf, flags, epoch, err, xq);
        else
            return (int)hipErrorInvalidValue;
    } else {
        if constexpr (G == 1 || G == 2)
            LK_LAUNCH((k_bwd_sn<T, NW, G>), G > 1 ? grid : ntask, 64 * NW, 0, stream, tasks, (const T*)panel,
                      (const T*)diag, rows, x, below, pbuf, flags, epoch, err, xq);
        else
            return (int)hipErrorInvalidValue;
    }
    return lk_last_error();
}

// ------------------------------------------------------------------------------------------
// Wide level solves, for levels of few supernodes with many blocks (a dense top separator),
// where a few workgroups per supernode would leave the device idle.  One 256-thread workgroup
// per row tile (forward) or per block column (backward) of every supernode of the level; the
// workgroups take their tasks from one ticket counter, in an order where every task's
// producers come first, so the earliest unfinished task can always run.  The tasks of one
// supernode run on all XCDs, so a solved block is handed over through the device-coherent
// level: the producer writes it with device-scope relaxed atomic stores (written through its
// XCD's L2), waits until they completed and sets the block's flag to the solve's epoch;
// consumers poll the flag and read the block with device-scope atomic loads.  Everything else
// a task reads is final before the launch (panel, diagonal copies, its own right-hand side, the
// ancestors' entries).  Wavefront v handles columns 16 v .. 16 v + 15 of each 64-column block
// (lane = row of a 64-row tile), so the four wavefronts follow the chain of blocks in step.
//  forward   task (s, t): acc = sum over blocks j < min(t, nb) of L(t, j) z_j, each block's
//            panel data loaded one block ahead of its wait; a diagonal tile solves
//            z_t = L_tt^{-1} (b_t - acc) and publishes it, a tile of rows below the supernode
//            adds -acc to the ancestors' entries (atomics).
//  backward  task (s, j): acc = sum over tiles i > j of L(i, j)^T x_i: the rows below the
//            supernode first (their x is final), then the diagonal tiles nb-1 .. j+1 as their
//            blocks are published; x_j = L_jj^{-T} (z_j - acc) is published.
// Per lane and column the products are summed in tile order, the wavefronts' or rows' partial
// sums in a fixed order: results do not depend on timing.
__device__ __forceinline__ bool wide_wait(const uint32_t* flag, uint32_t ep, int32_t* err) {
    long long n = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ep) {
        if ((n & 255) == 255 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (++n >= kSpinLimit) {
            if ((threadIdx.x & 63) == 0) atomicOr(err, 0x10);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}
__device__ __forceinline__ double wide_load(const double* p) {
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wide_store(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// after the wavefront's block stores: wait until they completed, then lane 0 sets the flag
__device__ __forceinline__ void wide_publish(uint32_t* flag, uint32_t ep) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if ((threadIdx.x & 63) == 0) __hip_atomic_store(flag, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// next ticket -> task index (-1: none left)
__device__ __forceinline__ int wide_task(const int32_t* order, int ntask, uint32_t* cnt, int* s_ti) {
    if (threadIdx.x == 0) {
        const int k = (int)atomicAdd(cnt, 1u);
        *s_ti = k < ntask ? order[k] : -1;
    }
    __syncthreads();
    return *s_ti;
}

// 16 columns (16 v + c of block j) of the 64-row tile starting at `row` (clamped to a panel row
// by the caller); columns >= kb and lanes with !rok read as zero (clamped unconditional loads)
template <typename T>
__device__ __forceinline__ void wide_cols(T (&h)[16], const T* __restrict__ P, int64_t ld, int j, int kb,
                                          int64_t row, bool rok) {
    const int v = threadIdx.x >> 6;
    const T* base = P + (int64_t)j * TS * ld + row;
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int cc = 16 * v + c;
        const T e = base[(int64_t)min(cc, kb - 1) * ld];
        h[c] = (rok && cc < kb) ? e : T(0);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fwd_wide(const lkd_task* __restrict__ tasks,
                                                  const int32_t* __restrict__ order, int ntask,
                                                  uint32_t* cnt, const T* __restrict__ L,
                                                  const T* __restrict__ diag,
                                                  const int32_t* __restrict__ sn_rows, double* x,
                                                  uint32_t* flags, const uint32_t* epoch,
                                                  int32_t* err) {
    __shared__ int s_ti, s_ok;
    __shared__ double Ds[TS * 65];
    __shared__ double red[4][TS];
    const int ti = wide_task(order, ntask, cnt, &s_ti);
    if (ti < 0) return;
    const lkd_task tk = tasks[ti];
    const uint32_t ep = *epoch;
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const Tiles tl(ld, tk.ncol);
    const int nb = tl.nb, nc = tk.ncol, t = tk.r0;
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t rb = tl.beg(t);
    const bool rok = rb + lane < tl.end(t);
    const int64_t row = rok ? rb + lane : rb;
    const int nj = min(t, nb);
    const uint32_t* fl = flags + tk.px;
    double* xs = x + tk.first;
    const int kbt = t < nb ? min(TS, nc - t * TS) : 0;
    if (threadIdx.x == 0) s_ok = 1;
    if (t < nb && v == 0) stage_diag<T, true>(diag + tk.aux + (int64_t)t * TS * TS, kbt, Ds);
    // the tile's right-hand side is final before the launch: loaded ahead of the chain
    const double b = lane < kbt ? xs[(int64_t)t * TS + lane] : 0.0;
    double acc = 0.0;
    bool ok = true;
    // Blocks are published in order (block j + 1 is solved after block j arrived), so a flag
    // seen set covers every block before it: blocks 0 .. known are published.  One flag wait
    // reads the next 63 flags too, and z comes four blocks per load: lane l of wavefront v holds
    // z of block zb + l / 16, column 16 v + l % 16 (blocks zb .. zmax).
    int known = -1, zb = 0, zmax = -1;
    double zg = 0.0;
    auto apply = [&](const T (&h)[16], int j) {
        if (!ok) return;
        if (j > zmax) {
            if (j > known) {
                if (!wide_wait(fl + j, ep, err)) {
                    ok = false;
                    return;
                }
                const int jj = j + 1 + lane;
                const bool set = jj < nj && __hip_atomic_load(fl + min(jj, nj - 1), __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT) == ep;
                const uint64_t m = __ballot(set);
                known = j + (~m == 0 ? 64 : __builtin_ctzll(~m));
            }
            zb = j;
            zmax = min(known, j + 3);
            const int jb = min(j + (lane >> 4), zmax);
            const int kb = min(TS, nc - jb * TS);
            const int cc = 16 * v + (lane & 15);
            const double zr = wide_load(xs + (int64_t)jb * TS + min(cc, kb - 1));
            zg = cc < kb ? zr : 0.0;
        }
        const int base = 16 * (j - zb);
#pragma unroll
        for (int c = 0; c < 16; ++c) acc = fma((double)h[c], bcast_lane(zg, base + c), acc);
    };
    // panel data two blocks ahead of the block applied
    T h0[16], h1[16], h2[16];
    auto ld_blk = [&](T (&h)[16], int j) { wide_cols<T>(h, P, ld, j, min(TS, nc - j * TS), row, rok); };
    if (nj > 0) ld_blk(h0, 0);
    if (nj > 1) ld_blk(h1, 1);
#pragma unroll 1
    for (int j = 0; j < nj; j += 3) {
        if (j + 2 < nj) ld_blk(h2, j + 2);
        apply(h0, j);
        if (j + 1 >= nj) break;
        if (j + 3 < nj) ld_blk(h0, j + 3);
        apply(h1, j + 1);
        if (j + 2 >= nj) break;
        if (j + 4 < nj) ld_blk(h1, j + 4);
        apply(h2, j + 2);
    }
    if (!ok) s_ok = 0;
    red[v][lane] = acc;
    __syncthreads();
    if (!s_ok || v != 0) return;
    const double s = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
    if (t < nb) {
        const double z = fsub_scaled(b - s, Ds, kbt);
        if (lane < kbt) wide_store(xs + (int64_t)t * TS + lane, z);
        wide_publish(flags + tk.px + t, ep);
    } else if (rok) {
        atomic_add(x + sn_rows[tk.aux2 + row], -s);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bwd_wide(const lkd_task* __restrict__ tasks,
                                                  const int32_t* __restrict__ order, int ntask,
                                                  uint32_t* cnt, const T* __restrict__ L,
                                                  const T* __restrict__ diag,
                                                  const int32_t* __restrict__ sn_rows, double* x,
                                                  uint32_t* flags, const uint32_t* epoch,
                                                  int32_t* err) {
    __shared__ int s_ti, s_ok;
    __shared__ double red[4][16][65];
    const int ti = wide_task(order, ntask, cnt, &s_ti);
    if (ti < 0) return;
    const lkd_task tk = tasks[ti];
    const uint32_t ep = *epoch;
    const T* P = L + tk.off;
    const int64_t ld = tk.ld;
    const Tiles tl(ld, tk.ncol);
    const int nb = tl.nb, nt = tl.nt, nc = tk.ncol, j = tk.r0;
    const int v = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kbj = min(TS, nc - j * TS);
    const uint32_t* fl = flags + tk.px + nb;
    double* xs = x + tk.first;
    if (threadIdx.x == 0) s_ok = 1;
    // the block's right-hand side z_j is final before the launch: loaded ahead of the chain
    const double rhs = lane < kbj ? xs[(int64_t)j * TS + lane] : 0.0;
    double part[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) part[c] = 0.0;
    bool ok = true;
    // tile i's 16 columns of block j, and (rows below the supernode) the ancestors' x
    auto load = [&](T (&h)[16], double& xr, int i) {
        const int64_t rb = tl.beg(i);
        const bool rok = rb + lane < tl.end(i);
        const int64_t row = rok ? rb + lane : rb;
        wide_cols<T>(h, P, ld, j, kbj, row, rok);
        if (i >= nb) xr = rok ? x[sn_rows[tk.aux2 + row]] : 0.0;
    };
    auto apply = [&](const T (&h)[16], double xr, int i) {
        if (!ok) return;
        if (i < nb) {
            if (!wide_wait(fl + i, ep, err)) {
                ok = false;
                return;
            }
            const int kb = min(TS, nc - i * TS);
            const double xv = wide_load(xs + (int64_t)i * TS + min(lane, kb - 1));
            xr = lane < kb ? xv : 0.0;
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) part[c] = fma((double)h[c], xr, part[c]);
    };
    // the diagonal block's columns (lane = column): row i of the row-major copy, diagonal
    // entry zeroed; rd = 1 / L_jj (stored on the copy's diagonal)
    T col[TS];
    double rd = 0.0;
    auto load_diag = [&]() {
        const T* Lr = diag + tk.aux + (int64_t)j * TS * TS;
        if (kbj == TS) {
#pragma unroll
            for (int i = 0; i < TS; ++i) col[i] = Lr[i * TS + lane];
        } else {
#pragma unroll
            for (int i = 0; i < TS; ++i)       // clamped loads times a 0/1 mask: no branches
                col[i] = Lr[min(i, kbj - 1) * kbj + min(lane, kbj - 1)] * ((i < kbj && lane < kbj) ? T(1) : T(0));
        }
#pragma unroll
        for (int i = 0; i < TS; ++i) col[i] = i == lane ? T(0) : col[i];
        const int dl = min(lane, kbj - 1);
        rd = lane < kbj ? (double)Lr[dl * kbj + dl] : 0.0;
    };
    T ha[16], hb[16];
    double xa = 0.0, xb = 0.0;
    int i = nt - 1;
    if (i > j) load(ha, xa, i);
    // ha holds tile i; the last tile, j + 1, is applied after the diagonal copy is loaded
#pragma unroll 1
    while (i > j + 1) {
        load(hb, xb, i - 1);
        apply(ha, xa, i);
        --i;
        if (i == j + 1) break;
        load(ha, xa, i - 1);
        apply(hb, xb, i);
        --i;
    }
    if (v == 0) load_diag();
    if (i == j + 1) {
        if (((nt - 1 - i) & 1) == 0) apply(ha, xa, i);
        else apply(hb, xb, i);
    }
    if (!ok) s_ok = 0;
#pragma unroll
    for (int c = 0; c < 16; ++c) red[v][c][lane] = part[c];
    __syncthreads();
    if (!s_ok || v != 0) return;
    // lane l: column l of the block = column l & 15 of wavefront l >> 4, rows in order
    const double* rr = red[lane >> 4][lane & 15];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int r = 0; r < TS; r += 4) {
        s0 += rr[r];
        s1 += rr[r + 1];
        s2 += rr[r + 2];
        s3 += rr[r + 3];
    }
    const double s = (s0 + s1) + (s2 + s3);
    const double val = lane < kbj ? rhs - s : 0.0;
    const double sol = bsub_regs<T>(val, col, rd);
    if (lane < kbj) wide_store(xs + (int64_t)j * TS + lane, sol);
    wide_publish(flags + tk.px + nb + j, ep);
}

// a new solve: the epoch advances and the queue counters restart
__global__ void k_epoch(uint32_t* epoch, uint32_t* cnt, int ncnt) {
    if (threadIdx.x == 0) atomicAdd(epoch, 1u);
    for (int i = threadIdx.x; i < ncnt; i += blockDim.x) cnt[i] = 0u;
}

// wavefronts per workgroup 4 or 8, workgroups per supernode 1 or 4 (forward, 8 wavefronts; the
// backward with 4 measured slower than with 1 and is not offered)
template <typename T>
int level_solve_cfg(int nw, int g, int dir, int32_t ntask, const lkd_task* tasks, const void* panel,
                    const void* diag, const int32_t* rows, double* x, double* below, double* pbuf,
                    uint32_t* flags, const uint32_t* epoch, int32_t* err, const XcdQueue& xq,
                    int grid, void* stream) {
    if (nw == 8 && g == 2)
        return level_solve<T, 8, 2>(dir, ntask, tasks, panel, diag, rows, x, below, pbuf, flags, epoch, err, xq, grid, stream);
    if (nw == 8 && g == 4)
        return level_solve<T, 8, 4>(dir, ntask, tasks, panel, diag, rows, x, below, pbuf, flags, epoch, err, xq, grid, stream);
    if (nw == 8)
        return level_solve<T, 8, 1>(dir, ntask, tasks, panel, diag, rows, x, below, pbuf, flags, epoch, err, xq, grid, stream);
    return level_solve<T, 4, 1>(dir, ntask, tasks, panel, diag, rows, x, below, pbuf, flags, epoch, err, xq, grid, stream);
}

}  // namespace

extern "C" {

int lkd_trsv_level(int prec, int dir, int nw, int g, int32_t ntask, const lkd_task* tasks,
                   const void* panel, const void* diag, const int32_t* sn_rows, double* x,
                   double* below, double* pbuf, uint32_t* flags, const uint32_t* epoch,
                   const int32_t* queue, const int32_t* qoff, int32_t nx, uint32_t* cnt,
                   int32_t grid, int32_t* err, void* stream) {
    if (ntask <= 0) return 0;
    if ((nw != 4 && nw != 8) || (g != 1 && !(g == 4 && nw == 8 && dir == 0) && !(g == 2 && nw == 8 && dir == 1)))
        return (int)hipErrorInvalidValue;
    if (g > 1 && (!queue || !qoff || !cnt || nx < 1 || grid < 1)) return (int)hipErrorInvalidValue;
    const XcdQueue xq{queue, qoff, cnt, nx};
    return prec == LKD_F32
               ? level_solve_cfg<float>(nw, g, dir, ntask, tasks, panel, diag, sn_rows, x, below, pbuf, flags, epoch, err, xq, grid, stream)
               : level_solve_cfg<double>(nw, g, dir, ntask, tasks, panel, diag, sn_rows, x, below, pbuf, flags, epoch, err, xq, grid, stream);
}

int lkd_trsv_wide(int prec, int dir, int32_t ntask, const lkd_task* tasks, const int32_t* order,
                  uint32_t* cnt, const void* panel, const void* diag, const int32_t* sn_rows,
                  double* x, uint32_t* flags, const uint32_t* epoch, int32_t* err, void* stream) {
    if (ntask <= 0) return 0;
    if ((dir != 0 && dir != 1) || !tasks || !order || !cnt || !flags || !epoch) return (int)hipErrorInvalidValue;
    if (prec == LKD_F32) {
        if (dir == 0)
            LK_LAUNCH(k_fwd_wide<float>, ntask, 256, 0, stream, tasks, order, ntask, cnt, (const float*)panel,
                      (const float*)diag, sn_rows, x, flags, epoch, err);
        else
            LK_LAUNCH(k_bwd_wide<float>, ntask, 256, 0, stream, tasks, order, ntask, cnt, (const float*)panel,
                      (const float*)diag, sn_rows, x, flags, epoch, err);
    } else {
        if (dir == 0)
            LK_LAUNCH(k_fwd_wide<double>, ntask, 256, 0, stream, tasks, order, ntask, cnt, (const double*)panel,
                      (const double*)diag, sn_rows, x, flags, epoch, err);
        else
            LK_LAUNCH(k_bwd_wide<double>, ntask, 256, 0, stream, tasks, order, ntask, cnt, (const double*)panel,
                      (const double*)diag, sn_rows, x, flags, epoch, err);
    }
    return lk_last_error();
}

int lkd_trsv_epoch(uint32_t* epoch, uint32_t* cnt, int32_t ncnt, void* stream) {
    if (ncnt < 0) return (int)hipErrorInvalidValue;
    LK_LAUNCH(k_epoch, 1, 256, 0, stream, epoch, cnt, ncnt);
    return lk_last_error();
}

}  // extern "C"
