"""1+1D adaptive meshes, spaces, saddle system and X indicators (oracle; test
infrastructure only).

Integer part: a Python restatement of the reference's 1+1D cell tree and DoF numbering,
written from the reference, independent of the library's C++ one:
* Tree.refine: space-time split into (left,bottom), (right,bottom), (left,top),
  (right,top) appended in that order, midpoint vertices through a registry keyed by
  llround(1e12 x) (mesh/refinement/refine_1d.hpp:221-310, split_policy.hpp:8-26,
  mesh/detail/vertex_registry.hpp, core/coord_key.hpp);
* Space.adjacency: faces grouped by fixed vertex, groups visited in the iteration order of
  a std::unordered_map<int, ...> reserved for 2 x (active cells) (taken from
  hash_order.cpp), exact matches first, then the nearest containing open face
  (fespace/adjacency/adjacency_1d.hpp:126-394);
* Space.refine: waves of splits with one-level grading closure
  (fespace/refinement/refine_1d.hpp:79-321), active list updated by swap-removal and
  sorted appends (fespace/partition_view.hpp:197-311, 1014-1045);
* Space.distribute: dof_distribution_1d.hpp:100-671 (elimination, slaves, unions,
  vertex unification, first-occurrence numbering, constrained DoFs);
* Space.restriction: assembly/constrained_dofs.hpp:133-226.
Pinned by: the published X / Y true-DoF counts of the uniform runs (16 x 16: 255 / 480,
algorithm_data_plot/1d/smooth_initial.csv:442) and of the adaptive runs (p = 1, 2, 3,
the X sequence 3, 5, 8, 12, 19, 31, 48, 80, ... of smooth_initial.csv:336-343 and its
p = 2, 3 counterparts), which the oracle's own adaptive chain reproduces; and by the
continuity of every X function across every face and of every Y function across spatial
faces (checked in tests from the restriction rows).

Floating-point part: element-local integrals on each Y cell with the X basis of its X
ancestor evaluated at the mapped Gauss points (the forms of
dense_local_main_blocks_2d.hpp:176-361 in 1+1D: A_y = int M phi_x phi_x,
B = int (psi_t phi + M psi_x phi_x), C = -int_{t=0} psi psi, g = -int_{t=0} u0 psi),
scattered through the restriction rows; a sparse direct solve; the X indicators
lambda_Y^2 + initial-trace^2 (indicator_utilities.hpp:109-171, 333-396, 435-473).
Pinned by: the published uniform-run totals (tests/golden/anchors.json) and the
published adaptive-run totals eta^2, lambda_Y^2 (tests/golden/adaptive1d_anchors.json).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .basis import segment_lagrange
from .marking import doerfler_mark, y_estimator_threshold
from .quadrature import gauss_legendre_01

HERE = os.path.dirname(os.path.abspath(__file__))
_HASH_SRC = os.path.join(HERE, "hash_order.cpp")
_HASH_LIB = os.path.join(HERE, "libapf_oracle_hash.so")
_hash = None


def build():
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-o", _HASH_LIB + ".tmp", _HASH_SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hash-order helper build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(_HASH_LIB + ".tmp", _HASH_LIB)
    return _HASH_LIB


def _hash_order(reserve_n, keys):
    """distinct keys in std::unordered_map<int, ...> iteration order"""
    global _hash
    if _hash is None:
        if not os.path.exists(_HASH_LIB):
            build()
        _hash = ctypes.CDLL(_HASH_LIB)
        _hash.oracle_hash_order.restype = ctypes.c_int64
        _hash.oracle_hash_order.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    k = np.ascontiguousarray(keys, dtype=np.int32)
    out = np.zeros(max(len(k), 1), dtype=np.int32)
    n = _hash.oracle_hash_order(reserve_n, len(k), k.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
    return [int(v) for v in out[:n]]


def _key(x):
    """std::llround(1e12 x): nearest integer, halves away from zero"""
    v = 1e12 * x
    return int(np.sign(v) * np.floor(abs(v) + 0.5))


class Tree:
    """cells: dicts with parent, gen, leaf, sv [x-left id, x-right id], tv [t0 id, t1 id], children"""

    def __init__(self, x0=0.0, x1=1.0, t0=0.0, t1=1.0):
        self.xv, self.tv, self.xreg, self.treg = [], [], {}, {}
        sv = [self._v(self.xv, self.xreg, x0), self._v(self.xv, self.xreg, x1)]
        tv = [self._v(self.tv, self.treg, t0), self._v(self.tv, self.treg, t1)]
        self.xb, self.tb = set(sv), set(tv)
        self.cells = [dict(parent=-1, gen=0, leaf=True, sv=sv, tv=tv, children=[])]

    @staticmethod
    def _v(V, reg, x):
        k = _key(x)
        if k not in reg:
            reg[k] = len(V)
            V.append(x)
        return reg[k]

    def refine(self, c):
        P = self.cells[c]
        assert P["leaf"]
        xm = self._v(self.xv, self.xreg, 0.5 * (self.xv[P["sv"][0]] + self.xv[P["sv"][1]]))
        tm = self._v(self.tv, self.treg, 0.5 * (self.tv[P["tv"][0]] + self.tv[P["tv"][1]]))
        xs = [[P["sv"][0], xm], [xm, P["sv"][1]]]
        ts = [[P["tv"][0], tm], [tm, P["tv"][1]]]
        kids = []
        for it in range(2):
            for ix in range(2):
                kids.append(len(self.cells))
                self.cells.append(dict(parent=c, gen=P["gen"] + 1, leaf=True, sv=list(xs[ix]), tv=list(ts[it]),
                                       children=[]))
        P["leaf"] = False
        P["children"] = kids

    def box(self, c):
        C = self.cells[c]
        return (self.xv[C["sv"][0]], self.xv[C["sv"][1]], self.tv[C["tv"][0]], self.tv[C["tv"][1]])

    def leaves(self):
        return [i for i, c in enumerate(self.cells) if c["leaf"]]


class Space:
    """X (ct=True) or Y (ct=False) space of degree p on a Tree."""

    def __init__(self, tree: Tree, p: int, ct: bool, active):
        self.tree, self.p, self.ct = tree, p, ct
        self.active = list(active)
        self.adjacency()
        self.distribute()

    # local dof l = i * (p+1) + j: i spatial node, j time node
    def sfd(self, face, j):
        return (0 if face == 0 else self.p) * (self.p + 1) + j

    def tfd(self, face, i):
        return i * (self.p + 1) + (0 if face == 0 else self.p)

    def adjacency(self):
        T = self.tree
        act = set(self.active)
        self.sif, self.tif = [], []
        for spatial in (True, False):
            recs = []
            for c in self.active:
                C = T.cells[c]
                x0, x1, t0, t1 = T.box(c)
                for face in (0, 1):
                    if spatial:
                        recs.append((C["sv"][face], c, face, C["tv"][0], C["tv"][1], t0, t1))
                    else:
                        recs.append((C["tv"][face], c, face, C["sv"][0], C["sv"][1], x0, x1))
            groups = {}
            for r in recs:
                groups.setdefault(r[0], []).append(r)
            order = _hash_order(2 * len(self.active), [r[0] for r in recs])
            out = self.sif if spatial else self.tif
            for key in order:
                g = sorted(groups[key], key=lambda r: (r[5], -r[6], r[1], r[2]))
                matched = [False] * len(g)
                stack = []
                for i, cur in enumerate(g):
                    while stack and g[stack[-1]][6] <= cur[5]:
                        stack.pop()
                    found = False
                    for k in reversed(stack):
                        cand = g[k]
                        if (cand[3], cand[4]) == (cur[3], cur[4]):
                            if cand[1] < cur[1]:
                                out.append(dict(mc=cand[1], mf=cand[2], sc=cur[1], sf=cur[2], hanging=False,
                                                boundary=False))
                            matched[i] = matched[k] = True
                            found = True
                            break
                    if not found:
                        for k in reversed(stack):
                            cand = g[k]
                            if cand[5] <= cur[5] and cur[6] <= cand[6]:
                                out.append(dict(mc=cand[1], mf=cand[2], sc=cur[1], sf=cur[2], hanging=True,
                                                boundary=False))
                                matched[i] = matched[k] = True
                                break
                    stack.append(i)
                for i, r in enumerate(g):
                    if matched[i]:
                        continue
                    C = T.cells[r[1]]
                    on_bnd = (C["sv"][r[2]] in T.xb) if spatial else (C["tv"][r[2]] in T.tb)
                    if on_bnd:
                        out.append(dict(mc=r[1], mf=r[2], sc=-1, sf=-1, hanging=False, boundary=True))
        assert all(f["mc"] in act for f in self.sif + self.tif)

    def _update_active(self, add, rem):
        pos = {c: i for i, c in enumerate(self.active)}
        for c in sorted(rem):
            if c not in pos:
                continue
            at = pos.pop(c)
            last = self.active.pop()
            if last != c:
                self.active[at] = last
                pos[last] = at
        for c in sorted(add):
            if c not in pos:
                pos[c] = len(self.active)
                self.active.append(c)

    def _wave(self, marked):
        if not marked:
            return False
        T = self.tree
        act = set(self.active)
        add, rem, children = set(), set(), []
        for c in marked:
            if c not in act:
                continue
            if T.cells[c]["leaf"]:
                T.refine(c)
            rem.add(c)
            for ch in T.cells[c]["children"]:
                if ch not in add:
                    children.append(ch)
                add.add(ch)
        if not children:
            return False
        self._update_active(add, rem)
        self.adjacency()
        act = set(self.active)
        forced, seen = [], set()
        for ifs in (self.tif, self.sif):  # level checks; in 1+1D both levels equal the generation
            for f in ifs:
                if f["boundary"] or not f["hanging"] or f["sc"] < 0:
                    continue
                if f["mc"] not in act or f["sc"] not in act:
                    continue
                if T.cells[f["sc"]]["gen"] > T.cells[f["mc"]]["gen"] + 1 and f["mc"] not in seen:
                    seen.add(f["mc"])
                    forced.append(f["mc"])
        self._wave(forced)
        return True

    def refine(self, marked):
        act, uniq, seen = set(self.active), [], set()
        for c in marked:
            c = int(c)
            if c in act and c not in seen:
                seen.add(c)
                uniq.append(c)
        if not uniq:
            return
        if not self._wave(uniq):
            self.adjacency()
        self.distribute()

    def coord(self, c, l):
        x0, x1, t0, t1 = self.tree.box(c)
        i, j = divmod(l, self.p + 1)
        return x0 + (x1 - x0) * (i / self.p), t0 + (t1 - t0) * (j / self.p)

    def distribute(self):
        T, p = self.tree, self.p
        NL = (p + 1) ** 2
        pos = {c: a for a, c in enumerate(self.active)}
        n = len(self.active) * NL
        lin = lambda c, l: pos[c] * NL + l
        elim = np.zeros(n, bool)
        for a, c in enumerate(self.active):
            C = T.cells[c]
            for l in range(NL):
                i = l // (p + 1)
                for face, at in ((0, 0), (1, p)):
                    if i == at and C["sv"][face] in T.xb:
                        elim[a * NL + l] = True
        slave = [None] * n
        lists = [(True, self.sif)] + ([(False, self.tif)] if self.ct else [])
        for spatial, ifs in lists:
            for k, f in enumerate(ifs):
                if not f["hanging"]:
                    continue
                for m in range(p + 1):
                    l = self.sfd(f["sf"], m) if spatial else self.tfd(f["sf"], m)
                    idx = lin(f["sc"], l)
                    if not elim[idx] and slave[idx] is None:
                        slave[idx] = (spatial, k)
        parent = list(range(n))

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a

        def unite(a, b):
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)

        for spatial, ifs in lists:
            for f in ifs:
                if f["boundary"] or f["hanging"]:
                    continue
                for m in range(p + 1):
                    la = self.sfd(f["mf"], m) if spatial else self.tfd(f["mf"], m)
                    lb = self.sfd(f["sf"], m) if spatial else self.tfd(f["sf"], m)
                    a, b = lin(f["mc"], la), lin(f["sc"], lb)
                    if not (elim[a] or elim[b] or slave[a] is not None or slave[b] is not None):
                        unite(a, b)
        if self.ct:
            owner = {}
            for idx in range(n):
                if elim[idx] or slave[idx] is not None:
                    continue
                l = idx % NL
                i, j = divmod(l, p + 1)
                if i not in (0, p) or j not in (0, p):
                    continue
                x, t = self.coord(self.active[idx // NL], l)
                key = (_key(x), _key(t))
                if key in owner:
                    unite(idx, owner[key])
                else:
                    owner[key] = idx
        gid_of_root = {}
        cd = -np.ones(n, np.int64)
        for idx in range(n):
            if elim[idx] or slave[idx] is not None:
                continue
            r = find(idx)
            if r not in gid_of_root:
                gid_of_root[r] = len(gid_of_root)
            cd[idx] = gid_of_root[r]
        self.n_true = len(gid_of_root)
        nodes = [i / p for i in range(p + 1)]
        bw = []
        for i in range(p + 1):
            prod = 1.0
            for j in range(p + 1):
                if j != i:
                    prod *= nodes[i] - nodes[j]
            bw.append(1.0 / prod)

        def interp(s):
            # barycentric second form with the node hit test (segment_lagrange.hpp:60-136):
            # t_i = w_i * (1 / (s - x_i)), L_i = t_i * (1 / sum t)
            for i in range(p + 1):
                if abs(s - nodes[i]) <= 32 * np.finfo(float).eps * max(1.0, abs(s) + abs(nodes[i])):
                    e = [0.0] * (p + 1)
                    e[i] = 1.0
                    return e
            t = [bw[i] * (1.0 / (s - nodes[i])) for i in range(p + 1)]
            s1 = 0.0
            for v in t:
                s1 += v
            inv = 1.0 / s1
            return [v * inv for v in t]

        def expand(c, l, factor, out):
            if abs(factor) < 1e-14:
                return
            idx = lin(c, l)
            if elim[idx]:
                return
            if slave[idx] is None:
                g = gid_of_root[find(idx)]
                out[g] = out.get(g, 0.0) + factor
                return
            spatial, k = slave[idx]
            x, t = self.coord(c, l)
            if spatial:
                f = self.sif[k]
                _, _, m0, m1 = T.box(f["mc"])
                w = interp((t - m0) / (m1 - m0))
                for m in range(p + 1):
                    expand(f["mc"], self.sfd(f["mf"], m), factor * w[m], out)
            else:
                f = self.tif[k]
                m0, m1, _, _ = T.box(f["mc"])
                w = interp((x - m0) / (m1 - m0))
                for m in range(p + 1):
                    expand(f["mc"], self.tfd(f["mf"], m), factor * w[m], out)

        slaves = [i for i in range(n) if not elim[i] and slave[i] is not None]
        slaves.sort(key=lambda i: (T.cells[self.active[i // NL]]["gen"], self.active[i // NL], i % NL))
        self.constraints = []  # (masters, weights) of constrained DoF n_true + k
        by_coord = {}
        for idx in slaves:
            c, l = self.active[idx // NL], idx % NL
            out = {}
            expand(c, l, 1.0, out)
            pairs = sorted((g, w) for g, w in out.items() if abs(w) > 1e-12)
            if not pairs:
                continue
            if len(pairs) == 1 and abs(pairs[0][1] - 1.0) < 1e-12:
                cd[idx] = pairs[0][0]
                continue
            x, t = self.coord(c, l)
            key = (_key(x), _key(t))
            if key in by_coord:
                cd[idx] = by_coord[key]
                continue
            g = self.n_true + len(self.constraints)
            self.constraints.append(([q for q, _ in pairs], [w for _, w in pairs]))
            by_coord[key] = g
            cd[idx] = g
        assert np.all((cd >= 0) | elim)
        self.cell_dofs = cd.reshape(len(self.active), NL)

    def restriction(self):
        """(ptr, dof, weight) rows of the active cells' local DoFs"""
        ptr, dof, w = [0], [], []
        for g in self.cell_dofs.ravel():
            if g >= 0:
                if g < self.n_true:
                    dof.append(int(g))
                    w.append(1.0)
                else:
                    m, ww = self.constraints[g - self.n_true]
                    dof.extend(m)
                    w.extend(ww)
            ptr.append(len(dof))
        return np.array(ptr, np.int64), np.array(dof, np.int32), np.array(w)

    def ancestors_in(self, coarse: "Space"):
        pos = {c: a for a, c in enumerate(coarse.active)}
        out = []
        for c in self.active:
            while c not in pos:
                c = self.tree.cells[c]["parent"]
                assert c >= 0
            out.append(pos[c])
        return np.array(out, np.int64)

    def local_values(self, coef):
        """[n_active, (p+1)^2] local coefficients of the function with true-DoF vector coef"""
        ptr, dof, w = self.restriction()
        NL = (self.p + 1) ** 2
        vals = np.zeros(len(self.active) * NL)
        for r in range(len(vals)):
            vals[r] = float(np.dot(w[ptr[r]:ptr[r + 1]], coef[dof[ptr[r]:ptr[r + 1]]]))
        return vals.reshape(len(self.active), NL)


def _tensor_basis(p, s, t):
    """values, d/dx (reference), d/dt (reference) of the (p+1)^2 local basis at points (s, t)"""
    vs, ds = segment_lagrange(p, s)
    vt, dt = segment_lagrange(p, t)
    V = np.einsum("qi,qj->qij", vs, vt).reshape(len(s), -1)
    Dx = np.einsum("qi,qj->qij", ds, vt).reshape(len(s), -1)
    Dt = np.einsum("qi,qj->qij", vs, dt).reshape(len(s), -1)
    return V, Dx, Dt


def assemble(X: Space, Y: Space, u0, kappa=1.0, ell=0.0):
    """K = [[A_y, B^T], [B, C]] (CSR) and b = [f; g] on the X / Y pair, element by element."""
    p = X.p
    nq = 2 * p + 1
    qs, ws = gauss_legendre_01(nq)
    S, Tq = np.meshgrid(qs, qs, indexing="ij")
    S, Tq = S.ravel(), Tq.ravel()
    W = np.outer(ws, ws).ravel()
    VY, DxY, DtY = _tensor_basis(p, S, Tq)
    rX = X.restriction()
    rY = Y.restriction()
    nX, nY = X.n_true, Y.n_true
    NL = (p + 1) ** 2
    anc = Y.ancestors_in(X)
    rows, cols, vals = [], [], []
    b = np.zeros(nY + nX)

    def scatter(rr, cr, roff, coff, loc):
        (rp, rd, rw), (cp, cdd, cw) = rr, cr
        for i in range(loc.shape[0]):
            for j in range(loc.shape[1]):
                v = loc[i, j]
                if abs(v) <= 1e-15:
                    continue
                for a in range(rp[i], rp[i + 1]):
                    for c in range(cp[j], cp[j + 1]):
                        rows.append(roff + rd[a])
                        cols.append(coff + cdd[c])
                        vals.append(rw[a] * cw[c] * v)

    def sub(r, cell):
        ptr, dof, w = r
        lo = ptr[cell * NL:(cell + 1) * NL + 1]
        return (lo - lo[0], dof[lo[0]:lo[-1]], w[lo[0]:lo[-1]])

    for yc in range(len(Y.active)):
        x0, x1, t0, t1 = Y.tree.box(Y.active[yc])
        hx, ht = x1 - x0, t1 - t0
        xa = X.active[anc[yc]]
        X0, X1, T0, T1 = X.tree.box(xa)
        xs, ts = x0 + hx * S, t0 + ht * Tq
        VX, DxX, DtX = _tensor_basis(p, (xs - X0) / (X1 - X0), (ts - T0) / (T1 - T0))
        mu = W * hx * ht
        gY = DxY / hx
        gX = DxX / (X1 - X0)
        tX = DtX / (T1 - T0)
        Ay = np.einsum("q,qi,qj->ij", mu * kappa, gY, gY)
        B = np.einsum("q,qi,qj->ij", mu, tX, VY) + np.einsum("q,qi,qj->ij", mu * kappa, gX, gY)
        ry, rx = sub(rY, yc), sub(rX, anc[yc])
        scatter(ry, ry, 0, 0, Ay)
        scatter(rx, ry, nY, 0, B)
        scatter(ry, rx, 0, nY, B.T)
        if ell != 0.0:
            fy = np.einsum("q,qi->i", mu * ell, VY)
            for i in range(NL):
                for a in range(ry[0][i], ry[0][i + 1]):
                    b[ry[1][a]] += ry[2][a] * fy[i]
    vs, _ = segment_lagrange(p, qs)
    for xc in range(len(X.active)):
        X0, X1, T0, T1 = X.tree.box(X.active[xc])
        if T0 != 0.0:
            continue
        h = X1 - X0
        Cl = -np.einsum("q,qi,qj->ij", ws * h, vs, vs)
        gl = -np.einsum("q,q,qi->i", ws * h, u0(X0 + h * qs), vs)
        rx = sub(rX, xc)
        loc_idx = [i * (p + 1) for i in range(p + 1)]  # bottom time node
        C = np.zeros((NL, NL))
        C[np.ix_(loc_idx, loc_idx)] = Cl
        scatter(rx, rx, nY, nY, C)
        for k, i in enumerate(loc_idx):
            for a in range(rx[0][i], rx[0][i + 1]):
                b[nY + rx[1][a]] += rx[2][a] * gl[k]
    K = sp.coo_matrix((vals, (rows, cols)), shape=(nY + nX, nY + nX)).tocsr()
    K.sum_duplicates()
    return K, b


def solve(K, b):
    return spla.spsolve(K.tocsc(), b)


def x_indicators(X: Space, Y: Space, x, u0, kappa=1.0):
    """(lambda_Y^2 per Y cell, initial-trace^2 per X cell) on the reference's rules
    (2p+1 Gauss points in space and time, indicator_utilities.hpp:109-171, 333-396)"""
    p = X.p
    nq = 2 * p + 1
    qs, ws = gauss_legendre_01(nq)
    S, Tq = np.meshgrid(qs, qs, indexing="ij")
    W = np.outer(ws, ws).ravel()
    _, DxY, _ = _tensor_basis(p, S.ravel(), Tq.ravel())
    lamloc = Y.local_values(x[:Y.n_true])
    ul = X.local_values(x[Y.n_true:])
    lam2 = np.zeros(len(Y.active))
    for yc in range(len(Y.active)):
        x0, x1, t0, t1 = Y.tree.box(Y.active[yc])
        g = DxY @ lamloc[yc] / (x1 - x0)
        lam2[yc] = np.sum(W * (x1 - x0) * (t1 - t0) * kappa * g * g)
    vs, _ = segment_lagrange(p, qs)
    tr2 = np.zeros(len(X.active))
    for xc in range(len(X.active)):
        X0, X1, T0, T1 = X.tree.box(X.active[xc])
        if T0 != 0.0:
            continue
        ub = vs @ ul[xc][[i * (p + 1) for i in range(p + 1)]]
        d = u0(X0 + (X1 - X0) * qs) - ub
        tr2[xc] = np.sum(ws * (X1 - X0) * d * d)
    return lam2, tr2


def adaptive_x_run(p, steps, theta_x=0.4, x_uniform_levels=1, u0=lambda x: np.sin(np.pi * x)):
    """Outer X loop (adaptive_driver.hpp:130-1230, with Y = X's mesh every outer
    iteration): per step the record (x_true_dofs, y_true_dofs, eta2, lambda_y2, trace2)
    and the marked X cells."""
    T = Tree()
    X = Space(T, p, True, T.leaves())
    for _ in range(x_uniform_levels):
        X.refine(list(X.active))
    out = []
    for _ in range(steps):
        Y = Space(T, p, False, X.active)
        K, b = assemble(X, Y, u0)
        x = solve(K, b)
        lam2_y, tr2_x = x_indicators(X, Y, x, u0)
        lam2_x = np.zeros(len(X.active))
        np.add.at(lam2_x, Y.ancestors_in(X), lam2_y)
        eta2 = lam2_x + tr2_x
        marked, _ = doerfler_mark(eta2, theta_x, ids=X.active)
        out.append(dict(x_true_dofs=X.n_true, y_true_dofs=Y.n_true, eta2=float(eta2.sum()),
                        lambda_y2=float(lam2_y.sum()), trace2=float(tr2_x.sum()), marked=list(marked)))
        X.refine(marked)
    return out


def adaptive_run(p, outer_steps, theta_x=0.4, theta_y=0.4, rho=1.0, max_inner=50, x_uniform_levels=1,
                 u0=lambda x: np.sin(np.pi * x)):
    """The adaptive driver with its inner Y loop (adaptive_driver.hpp:130-1230, the
    parameters of the published adaptive_standard runs: rho = 1, theta_x = theta_y = 0.4,
    deterministic reductions, no local time-slab closure, stop on an empty Y marking).  Per
    inner iteration: solve on (X, Y), X indicators, the Y estimator by source cell (flux^2 +
    reconstruction^2), its threshold rho^2 (lambda_Y^2 + (1 + 1/(rho + 1/2)^2) trace^2); accept
    when the estimator is below it, else Doerfler-mark the Y cells with theta_y and refine Y
    with its grading closure (time_slab_adaptive_refinement.hpp:240-338, closure disabled).
    After the accepted iteration the X cells are marked with theta_x on that iteration's
    indicators, X is refined and Y re-initialised on X's active cells (adaptive_driver.hpp:1129).
    Returns one record per inner iteration."""
    T = Tree()
    X = Space(T, p, True, T.leaves())
    for _ in range(x_uniform_levels):
        X.refine(list(X.active))
    Y = Space(T, p, False, X.active)
    out = []
    for outer in range(outer_steps):
        for inner in range(max_inner):
            K, b = assemble(X, Y, u0)
            x = solve(K, b)
            lam2_y, tr2_x = x_indicators(X, Y, x, u0)
            lam2_x = np.zeros(len(X.active))
            np.add.at(lam2_x, Y.ancestors_in(X), lam2_y)
            eta2 = lam2_x + tr2_x
            est = estimate_y(X, Y, x, u0)
            per_cell = est["flux_by_cell"] + est["recon_by_cell"]
            y2 = float(per_cell.sum())
            thr = y_estimator_threshold(rho, float(lam2_y.sum()), float(tr2_x.sum()))
            rec = dict(outer=outer, inner=inner, x_true_dofs=X.n_true, y_true_dofs_before=Y.n_true,
                       eta2=float(eta2.sum()), lambda_y2=float(lam2_y.sum()), trace2=float(tr2_x.sum()),
                       y_est2=y2, threshold2=thr, accepted=y2 <= thr)
            out.append(rec)
            if rec["accepted"]:
                rec["y_true_dofs_after"] = Y.n_true
                break
            marked_y, _ = doerfler_mark(per_cell, theta_y, ids=Y.active)
            rec["marked_y"] = list(marked_y)
            if not marked_y or inner + 1 >= max_inner:
                rec["y_true_dofs_after"] = Y.n_true
                break
            Y.refine(list(marked_y))
            rec["y_true_dofs_after"] = Y.n_true
        marked, _ = doerfler_mark(eta2, theta_x, ids=X.active)
        X.refine(marked)
        Y = Space(T, p, False, X.active)
    return out


def prolong_nodal(src: Space, coef, dst: Space):
    """true coefficients in `dst` of the `src` function: its value at the node of each free
    dst DoF, on the first active dst cell (active order, local order) where the DoF is free,
    evaluated on the src cell containing it (fespace/prolongation.hpp:15-97)"""
    anc = dst.ancestors_in(src)
    loc = src.local_values(coef)
    q = dst.p
    out = np.full(dst.n_true, np.nan)
    for cpos, c in enumerate(dst.active):
        x0, x1, t0, t1 = dst.tree.box(c)
        for l, g in enumerate(dst.cell_dofs[cpos]):
            if 0 <= g < dst.n_true and np.isnan(out[g]):
                i, j = divmod(l, q + 1)
                a = anc[cpos]
                out[g] = _value_on_cell(src, loc[a], a, x0 + (x1 - x0) * i / q, t0 + (t1 - t0) * j / q)
    assert not np.isnan(out).any()
    return out


def g_estimate(X: Space, Y: Space, x, u0, kappa=1.0):
    """G^2 of the solution x = [lambda; u] on the adaptive pair
    (compute_g_estimator_for_accepted_inner_iteration_, adaptive_driver.hpp:2680-2905): G =
    Y's cells with degree p+1 in space and time, A_G lambda_G = f_G - B(X,G)^T u (sparse
    direct solve), G^2 = d^T A_G d for d = lambda_G - I lambda.  A_G, f_G and B come from the
    degree-(p+1) assembly on (X_{p+1}, G), B applied to the prolongation I u."""
    p = X.p
    Xq = Space(X.tree, p + 1, True, X.active)
    G = Space(X.tree, p + 1, False, Y.active)
    K, b = assemble(Xq, G, u0, kappa)
    nG = G.n_true
    K = sp.csr_matrix(K)
    A, Bt = K[:nG, :nG], K[:nG, nG:]
    lam, u = x[:Y.n_true], x[Y.n_true:]
    Iu = prolong_nodal(X, u, Xq)
    Il = prolong_nodal(Y, lam, G)
    lg = spla.spsolve(A.tocsc(), b[:nG] - Bt @ Iu)
    d = lg - Il
    return dict(g_true_dofs=nG, g_lambda_difference_squared=max(0.0, float(d @ (A @ d))))


def graph_norm(X: Space, Y: Space, K, kappa=1.0, parts=False):
    """Hhat = -C_signed + B_dt diag(A_y)^-1 B_dt^T + A_x (assemble_parabolic_graph_norm_
    approximation, assembly/main_system/parabolic_graph_norm.hpp:22-65, with
    schur_approximations.hpp:80-287): B_dt = int dpsi/dt phi on the Y cells (X basis of the
    ancestor, mat_B.hpp:250-273), A_x = int kappa psi_x psi_x on the X cells with their own
    rule (mat_A.hpp), C_signed and A_y taken from K."""
    p = X.p
    nq = 2 * p + 1
    qs, ws = gauss_legendre_01(nq)
    S, Tq = np.meshgrid(qs, qs, indexing="ij")
    S, Tq = S.ravel(), Tq.ravel()
    W = np.outer(ws, ws).ravel()
    VY, _, _ = _tensor_basis(p, S, Tq)
    _, DxX0, _ = _tensor_basis(p, S, Tq)
    nX, nY = X.n_true, Y.n_true
    NL = (p + 1) ** 2
    anc = Y.ancestors_in(X)
    rX, rY = X.restriction(), Y.restriction()

    def expand(r, cell):
        ptr, dof, w = r
        P = np.zeros((NL, max(nX, nY)))
        for l in range(NL):
            for a in range(ptr[cell * NL + l], ptr[cell * NL + l + 1]):
                P[l, dof[a]] += w[a]
        return P

    Bdt = np.zeros((nX, nY))
    for yc in range(len(Y.active)):
        x0, x1, t0, t1 = Y.tree.box(Y.active[yc])
        X0, X1, T0, T1 = X.tree.box(X.active[anc[yc]])
        xs, ts = x0 + (x1 - x0) * S, t0 + (t1 - t0) * Tq
        _, _, DtX = _tensor_basis(p, (xs - X0) / (X1 - X0), (ts - T0) / (T1 - T0))
        loc = np.einsum("q,qi,qj->ij", W * (x1 - x0) * (t1 - t0), DtX / (T1 - T0), VY)
        Bdt += expand(rX, anc[yc])[:, :nX].T @ loc @ expand(rY, yc)[:, :nY]
    Ax = np.zeros((nX, nX))
    for xc in range(len(X.active)):
        X0, X1, T0, T1 = X.tree.box(X.active[xc])
        g = DxX0 / (X1 - X0)
        loc = np.einsum("q,qi,qj->ij", W * (X1 - X0) * (T1 - T0) * kappa, g, g)
        E = expand(rX, xc)[:, :nX]
        Ax += E.T @ loc @ E
    Kd = K.toarray()
    Ay_diag = np.diag(Kd[:nY, :nY])
    C = Kd[nY:, nY:]
    H = -C + (Bdt / Ay_diag[None, :]) @ Bdt.T + Ax
    return (H, Bdt, Ax) if parts else H


def _value_on_cell(space: Space, local, cell, x, t):
    """value at (x, t) of the function with local coefficients `local` [NL] on active cell
    position `cell` (Lagrange basis, local dof i_space * (p+1) + j_time)"""
    p = space.p
    x0, x1, t0, t1 = space.tree.box(space.active[cell])
    vs, _ = segment_lagrange(p, [(x - x0) / (x1 - x0)])
    vt, _ = segment_lagrange(p, [(t - t0) / (t1 - t0)])
    return float(np.outer(vs[0], vt[0]).ravel() @ local)


def slab_groups(Y: Space):
    """The time slabs between consecutive time endpoints of Y's cells
    (time_slabs/time_slab_builder.hpp:53-120, each source cell sliced into the slabs it
    spans), runs of consecutive slabs with one spatial partition grouped: per group the
    interval endpoints, the slab times and, per slab, the source cell (position in Y's
    active list) of each interval, left to right."""
    boxes = [Y.tree.box(c) for c in Y.active]
    times = sorted({b[2] for b in boxes} | {b[3] for b in boxes})
    groups = []
    for k in range(len(times) - 1):
        t0, t1 = times[k], times[k + 1]
        cells = sorted((b[0], b[1], a) for a, b in enumerate(boxes) if b[2] <= t0 and b[3] >= t1)
        part = tuple((c[0], c[1]) for c in cells)
        if groups and groups[-1]["part"] == part:
            groups[-1]["times"].append(t1)
            groups[-1]["src"].append([c[2] for c in cells])
        else:
            groups.append(dict(part=part, times=[t0, t1], src=[[c[2] for c in cells]]))
    return groups


def estimate_y(X: Space, Y: Space, x, u0, kappa=1.0, ell=0.0):
    """The Y estimator of x = [lambda; u] on the adaptive pair, slab group by slab group:
    on each group (slabs of one spatial partition = a tensor mesh) the tensor oracle
    estimator (oracle.estimator.estimate, per-slab direct reconstruction solves, per
    (slab, vertex) patches) on [lambda of the source cells at the group's Y nodes; u at its
    X nodes]; per-cell values summed into the source cells
    (time_slab_adaptive_loop.hpp:833-2026, by_source_cell)."""
    from . import assembly as OA
    from . import estimator as OE
    from . import mesh as OM

    p = Y.p
    Nt = p + 1
    lamloc = Y.local_values(x[:Y.n_true])
    uloc = X.local_values(x[Y.n_true:])
    xboxes = [X.tree.box(c) for c in X.active]
    anc = Y.ancestors_in(X)
    out = dict(flux2=0.0, recon2=0.0, flux_by_cell=np.zeros(len(Y.active)),
               recon_by_cell=np.zeros(len(Y.active)), resid_by_cell=np.zeros(len(Y.active)))
    for g in slab_groups(Y):
        verts = np.array([g["part"][0][0]] + [b for (_, b) in g["part"]])
        ns = len(g["part"])
        sm = OM.SpatialMesh(1, verts[:, None], np.stack([np.arange(ns), np.arange(1, ns + 1)], axis=1))
        stm = OM.SpaceTimeMesh(sm, np.array(g["times"]))
        od = OM.spacetime_dofs(stm, p)
        prob = OA.Problem(kappa * np.ones((ns, 1, 1)), ell, lambda pts: u0(pts[:, 0]))
        S = OA.assemble(stm, od, prob)
        m = stm.n_intervals
        lam = np.zeros(od.nY)
        u = np.zeros(od.nX)
        for k in range(m):
            tk, hk = stm.times[k], stm.times[k + 1] - stm.times[k]
            for c in range(ns):
                a = g["src"][k][c]
                xa, xb = verts[c], verts[c + 1]
                for i in range(p + 1):
                    xi = xa + (xb - xa) * i / p
                    for l in range(Nt):
                        t = tk + hk * l / p
                        yd = od.y_map[k * ns + c, i * Nt + l]
                        if yd >= 0:
                            lam[yd] = _value_on_cell(Y, lamloc[a], a, xi, t)
                        xd = od.x_map[k * ns + c, i * Nt + l]
                        if xd >= 0:
                            xa_cell = anc[a]
                            u[xd] = _value_on_cell(X, uloc[xa_cell], xa_cell, xi, t)
        r = OE.estimate(stm, od, prob, S, np.concatenate([lam, u]))
        out["flux2"] += r.flux2
        out["recon2"] += r.recon2
        for k in range(m):
            for c in range(ns):
                a = g["src"][k][c]
                for key in ("flux_by_cell", "recon_by_cell", "resid_by_cell"):
                    out[key][a] += getattr(r, key)[k * ns + c]
    out["y_est2"] = out["flux2"] + out["recon2"]
    del xboxes
    return out
