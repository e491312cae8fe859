"""CPU emulation of the device factorization / triangular-solve schedule (numpy).

Executes the exported task lists with the semantics of each kernel in
soverign-gpu-accelerator-_amd/csrc/hip/{chol,trsv,vec}.hip, so planning errors (wrong tiles,
levels, offsets, ancestor targets) show up on the CPU against a dense Cholesky."""
import numpy as np

TS = 64


class Emulator:
    def __init__(self, plan):
        self.p = plan
        self.fl = plan["flaunch"].reshape(-1, 3)
        self.sl = plan["slaunch"].reshape(-1, 3)
        self.panel = np.zeros(int(plan["sn_panel"][-1]))
        self.diag = np.zeros(max(int(plan["diag_elems"][0]), 1))
        self.piv = np.zeros(len(plan["perm"]))
        self.err = 0

    def blk(self, t, r0, r1, c0, c1):
        """view of panel rows r0:r1, cols c0:c1 of the task's supernode (column-major)."""
        ld = int(t["ld"])
        base = int(t["off"])
        M = self.panel[base:base + ld * int(t["ncol"])].reshape(int(t["ncol"]), ld).T
        return M[r0:r1, c0:c1]

    def assemble(self, theta, equil=False, delta=1e-9):
        p = self.p
        prod = p["pp"] * theta[p["pk"]]
        seg = np.add.reduceat(prod, p["segptr"][:-1]) if len(prod) else np.zeros(0)
        m = len(p["perm"])
        D = np.ones(m)
        if equil:
            mii = np.zeros(m)
            ok = p["dseg"] >= 0
            mii[ok] = seg[p["dseg"][ok]]
            D = 1.0 / np.sqrt(np.maximum(np.abs(mii), 1e-12))
        v = seg * D[p["segi"]] * D[p["segj"]]
        diag = (p["segi"] == p["segj"]) & (p["empty_row"][p["segi"]] == 0)
        v[diag] += delta
        self.panel[:] = 0
        self.panel[p["segdst"]] = v
        return D

    def _factor_block(self, A):
        try:
            return np.linalg.cholesky(A)
        except np.linalg.LinAlgError:
            self.err |= 1
            return None

    def _diag_copy(self, L, off):
        kb = L.shape[0]
        Lr = np.tril(L, -1).copy()
        Lr[np.arange(kb), np.arange(kb)] = 1.0 / np.diag(L)
        self.diag[off:off + kb * kb] = Lr.reshape(-1)

    def dag_order(self, b, c, rng=None):
        """An execution order of STEPS launch tasks b .. b+c that respects their dependencies:
        list order, or (rng given) a random choice among the ready tasks at every step."""
        if rng is None:
            return list(range(b, b + c))
        deps = self.p["fdeps"][4 * b:4 * (b + c)].reshape(c, 4)
        indeg = (deps >= 0).sum(axis=1)
        kids = [[] for _ in range(c)]
        for i in range(c):
            for d in deps[i]:
                if d >= 0:
                    kids[d].append(i)
        ready = [i for i in range(c) if indeg[i] == 0]
        out = []
        while ready:
            i = ready.pop(int(rng.integers(len(ready))))
            out.append(b + i)
            for j in kids[i]:
                indeg[j] -= 1
                if indeg[j] == 0:
                    ready.append(j)
        assert len(out) == c, "STEPS dependencies have a cycle"
        return out

    def factor(self, rng=None):
        """Left-looking schedule: DIAG (kind 1, block column 0), PART (6; with nsplit = 2 the
        step's diagonal block, as DIAG), PSUM (8: a tile's partial products summed in order
        into the first), STEP (5), STEPS (7:
        the STEP tiles of a level, run in list order or, with rng, in a random order that
        respects their dependencies), UPDATE (2), DIAGWB (3).
        Diagonal factors live in staging slots between launches; the panel's diagonal blocks
        are only written by DIAGWB."""
        p = self.p
        T = p["ftasks"]
        MT = 128
        self.stage = {}
        self.work = {}
        self.psum_slices = {}
        for kind, b, c in self.fl:
            assert kind in (1, 2, 3, 5, 6, 7, 8)
            order = self.dag_order(b, c, rng) if kind == 7 else range(b, b + c)
            steps = kind == 7
            for q in order:
                t = T[q]
                if steps:                         # STEPS: PART tasks and STEP tiles
                    kind = 6 if int(t["part"]) else 5
                if self.err:
                    return False
                k0, kb, r0, c0, ld, nc = (int(t[x]) for x in ("k0", "kb", "r0", "c0", "ld", "ncol"))
                first = int(t["first"])
                if kind == 1 or (kind == 6 and int(t["nsplit"]) == 2):
                    assert kind == 6 or k0 == 0
                    D = self.blk(t, k0, k0 + kb, k0, k0 + kb).copy()
                    L = self._factor_block(np.tril(D) + np.tril(D, -1).T)
                    if L is None:
                        return False
                    self._publish(L, int(t["ws"]), first + k0, int(t["aux"]))
                elif kind == 8:                   # PSUM: slice kbeg of the tile's np products
                    px, npt, sl = int(t["px"]), int(t["np"]), int(t["kbeg"])
                    assert 0 <= sl < 16 and npt >= 2
                    self.psum_slices.setdefault(px, set()).add(sl)
                    if len(self.psum_slices[px]) == 16:   # all slices: the products become one
                        got = [self.work.pop(px + s_ * 128 * 64) for s_ in range(npt)]
                        got.sort()
                        assert all(got[i][1] == got[i + 1][0] for i in range(npt - 1))
                        tot = got[0][2]
                        for g_ in got[1:]:
                            tot = tot + g_[2]
                        self.work[px] = (got[0][0], got[-1][1], tot)
                        del self.psum_slices[px]
                elif kind == 6:                   # PART: one K range of a STEP tile's GEMM
                    ra, kbeg, kend = int(t["c0"]), int(t["kbeg"]), int(t["kend"])
                    assert 0 <= kbeg < kend <= k0
                    px = int(t["px"])
                    assert px not in self.work
                    Ar = self.blk(t, r0, r0 + ra, kbeg, kend)
                    self.work[px] = (kbeg, kend, Ar @ self.blk(t, k0, k0 + kb, kbeg, kend).T)
                    if int(t["part"]) == 2:       # and of the next block column, into ws
                        ws, kb2 = int(t["ws"]), min(64, nc - k0 - 64)
                        assert ws not in self.work and kb2 > 0 and r0 >= k0 + 128
                        self.work[ws] = (kbeg, kend, Ar @ self.blk(t, k0 + 64, k0 + 64 + kb2, kbeg, kend).T)
                elif kind == 5:
                    ra = int(t["c0"])
                    X = self.blk(t, r0, r0 + ra, k0, k0 + kb)
                    kbeg, kend = int(t["kbeg"]), int(t["kend"])
                    cover = [(kbeg, kend)] if kend > kbeg else []
                    G = self.blk(t, r0, r0 + ra, kbeg, kend) @ self.blk(t, k0, k0 + kb, kbeg, kend).T
                    for s in range(int(t["np"])):  # partial products of the PART launch before
                        pb, pe, Ps = self.work.pop(int(t["px"]) + s * 128 * 64)
                        cover.append((pb, pe))
                        G = G + Ps
                    cover.sort()                  # the K ranges tile [0, k0) exactly
                    assert [x for rg in cover for x in rg][:1] in ([], [0]) and \
                        all(cover[i][1] == cover[i + 1][0] for i in range(len(cover) - 1)) and \
                        (cover[-1][1] if cover else 0) == k0, cover
                    X -= G
                    L = self.stage[int(t["ws"])]
                    X[:] = np.linalg.solve(L, X.T).T
                    for h in range(2):            # diagonal blocks inside the tile
                        d0 = r0 + 64 * h
                        if 64 * h >= ra or d0 >= nc:
                            break
                        db = min(64, ra - 64 * h, nc - d0)
                        Xd = self.blk(t, d0, d0 + db, k0, k0 + kb)
                        C = self.blk(t, d0, d0 + db, d0, d0 + db)
                        C -= np.tril(Xd @ Xd.T)
                    if int(t["nsplit"]):
                        D = self.blk(t, r0, r0 + ra, r0, r0 + ra).copy()
                        L1 = self._factor_block(np.tril(D) + np.tril(D, -1).T)
                        if L1 is None:
                            return False
                        self._publish(L1, int(t["aux2"]), first + r0, int(t["aux"]))
                elif kind == 3:
                    Lr = self.diag[int(t["aux"]):int(t["aux"]) + kb * kb].reshape(kb, kb)
                    Lk = np.tril(Lr, -1) + np.diag(self.piv[first + k0:first + k0 + kb])
                    Dp = self.blk(t, k0, k0 + kb, k0, k0 + kb)
                    Dp[:] = np.where(np.tri(kb, dtype=bool), Lk, Dp)
                else:
                    kbeg, kend = int(t["kbeg"]), int(t["kend"])
                    rows = p["sn_rows"][int(t["aux2"]):int(t["aux2"]) + ld]
                    ra, rb = min(MT, ld - r0), min(64, ld - c0)
                    U = self.blk(t, r0, r0 + ra, kbeg, kend) @ self.blk(t, c0, c0 + rb, kbeg, kend).T
                    for i in range(ra):
                        for j in range(rb):
                            if r0 + i < c0 + j:
                                continue
                            gi, gj = rows[r0 + i], rows[c0 + j]
                            s = p["col_to_sn"][gj]
                            f = p["sn_first"][s]
                            rp0, rp1 = p["sn_rowptr"][s], p["sn_rowptr"][s + 1]
                            trows = p["sn_rows"][rp0:rp1]
                            pos = np.searchsorted(trows, gi)
                            if pos >= len(trows) or trows[pos] != gi:
                                self.err |= 2
                                return False
                            self.panel[p["sn_panel"][s] + (gj - f) * (rp1 - rp0) + pos] -= U[i, j]
        assert not self.work, "partial products never added"
        assert not self.psum_slices, "partial sums with missing slices"
        return True

    def _publish(self, L, slot, col, dcopy):
        self.stage[slot] = L
        self.piv[col:col + L.shape[0]] = np.diag(L)
        self._diag_copy(L, dcopy)

    def dense_L(self):
        """L in the plan's (postordered) permutation as a dense lower triangle."""
        p = self.p
        m = len(p["perm"])
        L = np.zeros((m, m))
        for s in range(len(p["sn_first"]) - 1):
            f, l = p["sn_first"][s], p["sn_first"][s + 1]
            rows = p["sn_rows"][p["sn_rowptr"][s]:p["sn_rowptr"][s + 1]]
            ld = len(rows)
            blk = self.panel[p["sn_panel"][s]:p["sn_panel"][s] + ld * (l - f)].reshape(l - f, ld).T
            for c in range(l - f):
                for r in range(ld):
                    if rows[r] >= f + c:
                        L[rows[r], f + c] = blk[r, c]
        return L

    # ---------------------------------------------------------------- triangular solves
    def _diag_L(self, off, kb):
        Lr = self.diag[off:off + kb * kb].reshape(kb, kb)
        return np.tril(Lr, -1) + np.diag(1.0 / np.diag(Lr))

    def solve(self, x):
        """In-place L L^T solve of the permuted vector x with the level schedule (kinds 20 / 21:
        G workgroups per supernode of a level (task g = r0 of c0), all with the same wavefront
        count; kinds 22 / 23, the wide form: one task per row tile / block column); per
        supernode the blocks are processed in order, rows below go to the ancestors' entries
        (forward) or are read from them (backward).  Checks the plan's workgroup sets
        and its scratch, exchange and flag ranges (disjoint per supernode) on the way."""
        p = self.p
        T = p["stasks"]
        seen_flags = []
        for kind, b, c in self.sl:
            assert kind in (20, 21, 22, 23)
            tl = T[b:b + c]
            if kind >= 22:                          # wide: one task per tile / block column
                heads, seen = [], set()
                for t in tl:
                    nbk = (int(t["ncol"]) + TS - 1) // TS
                    ntk = nbk + (int(t["ld"]) - int(t["ncol"]) + TS - 1) // TS
                    key = (int(t["sn"]), int(t["r0"]))
                    assert key not in seen and 0 <= key[1] < (ntk if kind == 22 else nbk)
                    seen.add(key)
                    if int(t["sn"]) not in {int(h["sn"]) for h in heads}:
                        heads.append(t)
                for t in heads:
                    nbk = (int(t["ncol"]) + TS - 1) // TS
                    ntk = nbk + (int(t["ld"]) - int(t["ncol"]) + TS - 1) // TS
                    assert {k for s_, k in seen if s_ == int(t["sn"])} == set(range(ntk if kind == 22 else nbk))
                    seen_flags.append((int(t["px"]), int(t["px"]) + 2 * nbk))
            else:
                assert len({int(u["nsplit"]) for u in tl}) == 1 and int(tl[0]["nsplit"]) in (4, 8)
                for t in tl:                        # G workgroups per supernode: g = 0 .. G-1 once each
                    G = int(t["c0"])
                    assert 1 <= G and 0 <= int(t["r0"]) < G
                    gs = sorted(int(u["r0"]) for u in tl if u["sn"] == t["sn"])
                    assert gs == list(range(G))
                heads = [t for t in tl if int(t["r0"]) == 0]
            for t in (heads if kind < 22 else []):
                nbk = (int(t["ncol"]) + TS - 1) // TS
                G, nw = int(t["c0"]), int(t["nsplit"])
                for u in heads:
                    if u["sn"] != t["sn"]:
                        a0, a1 = int(t["ws"]), int(t["ws"]) + nbk * G * nw * TS
                        b0 = int(u["ws"])
                        b1 = b0 + (int(u["ncol"]) + TS - 1) // TS * int(u["c0"]) * int(u["nsplit"]) * TS
                        assert a1 <= b0 or b1 <= a0
                        a0, a1 = int(t["kbeg"]), int(t["kbeg"]) + int(t["ld"]) - int(t["ncol"])
                        b0, b1 = int(u["kbeg"]), int(u["kbeg"]) + int(u["ld"]) - int(u["ncol"])
                        assert a1 <= b0 or b1 <= a0 or a0 == a1 or b0 == b1
                seen_flags.append((int(t["px"]), int(t["px"]) + nbk + nbk * G))
            for t in heads:
                ld, nc, first = int(t["ld"]), int(t["ncol"]), int(t["first"])
                rows = p["sn_rows"][int(t["aux2"]):int(t["aux2"]) + ld]
                nb = (nc + TS - 1) // TS
                Lfull = self.blk(t, 0, ld, 0, nc)
                Ld = np.zeros((nc, nc))
                for j in range(nb):
                    kb = min(TS, nc - j * TS)
                    Ld[j * TS:j * TS + kb, j * TS:j * TS + kb] = self._diag_L(int(t["aux"]) + j * TS * TS, kb)
                    Ld[j * TS + kb:, j * TS:j * TS + kb] = Lfull[j * TS + kb:nc, j * TS:j * TS + kb]
                if kind in (20, 22):
                    z = np.linalg.solve(Ld, x[first:first + nc])
                    x[first:first + nc] = z
                    if ld > nc:
                        np.subtract.at(x, rows[nc:ld], Lfull[nc:ld] @ z)
                else:
                    y = x[first:first + nc] - (Lfull[nc:ld].T @ x[rows[nc:ld]] if ld > nc else 0.0)
                    x[first:first + nc] = np.linalg.solve(Ld.T, y)
        fl = sorted(set(seen_flags))                # flag areas of different supernodes disjoint
        assert all(fl[i][1] <= fl[i + 1][0] for i in range(len(fl) - 1))
        return x
