"""Binary snapshots in the reference's formats: headers, block sizes and the
meaning of every block, read back the way the reference's reader parses them
(python/binary_readers.py:42-63 versioned header, :90-205 mesh file; the
per-cell files as write_cellwise_error_binary.hpp:54-77 and
write_cellwise_error_components_binary.hpp:36-86 write them)."""
import numpy as np
import pytest

from doubleadapparabolicfem_amd import combine_estimator, meshes, snapshot


def _header(f, payload, version):
    h = np.fromfile(f, dtype=np.int32, count=payload + 1)
    assert h.size == payload + 1
    assert h[0] == version
    return h[1:]


def _block(f, dtype, count, shape=None):
    a = np.fromfile(f, dtype=dtype, count=count)
    assert a.size == count
    return a.reshape(shape) if shape is not None else a


def read_mesh(path):
    with open(path, "rb") as f:
        h = _header(f, 13, 1)
        nsv, ntv, ds, dt, nc, vsf, vtf, tp, ip, nfs, msn, nft, mtn = (int(v) for v in h)
        out = dict(header=h,
                   spatial_vertices=_block(f, np.float64, nsv * ds, (nsv, ds)),
                   temporal_vertices=_block(f, np.float64, ntv * dt, (ntv, dt)),
                   cell_ids=_block(f, np.int32, nc),
                   spatial_ids=_block(f, np.int32, nc * tp, (nc, tp)),
                   temporal_ids=_block(f, np.int32, nc * ip, (nc, ip)),
                   spatial_neigh=_block(f, np.int32, nc * nfs * msn, (nc, nfs, msn)),
                   temporal_neigh=_block(f, np.int32, nc * nft * mtn, (nc, nft, mtn)))
        assert np.fromfile(f, dtype=np.uint8).size == 0  # no trailing bytes
    return out


def test_format_versions_are_the_reference_ones():
    # binary_format_versions.hpp:10, :20, :21
    assert snapshot.MESH_BINARY_FORMAT_VERSION == 1
    assert snapshot.CELLWISE_ERROR_BINARY_FORMAT_VERSION == 1
    assert snapshot.CELLWISE_ERROR_COMPONENTS_BINARY_FORMAT_VERSION == 1


@pytest.mark.parametrize("case", ["uj2", "interval", "fan"])
def test_mesh_file_describes_the_space_time_mesh(tmp_path, case):
    if case == "uj2":
        m = meshes.union_jack_mesh(2, np.linspace(0.0, 1.0, 4))
    elif case == "interval":
        m = meshes.uniform_mesh(1, 2)
    else:
        m = meshes.fan_mesh(2, meshes.graded_times(1, 3), seed=5)
    path = tmp_path / "mesh.bin"
    snapshot.write_mesh_binary(m, path)
    d = read_mesh(path)
    dim = m.dim
    V = np.asarray(m.vertices).reshape(-1, dim)
    C = np.asarray(m.cells).reshape(-1, dim + 1)
    T = np.asarray(m.times)
    ns, mm = len(C), len(T) - 1
    nsv, ntv, ds, dt, nc, vsf, vtf, tp, ip, nfs, msn, nft, mtn = (int(v) for v in d["header"])
    assert (nsv, ntv, ds, dt, nc) == (len(V), mm + 1, dim, 1, ns * mm)
    assert (vsf, vtf, tp, ip, nfs, nft) == (dim, dim + 1, dim + 1, 2, dim + 1, 2)
    assert np.array_equal(d["spatial_vertices"], V)
    assert np.array_equal(d["temporal_vertices"][:, 0], T)
    assert np.array_equal(d["cell_ids"], np.arange(ns * mm))
    # cell k * ns + c: the vertices of spatial cell c over [t_k, t_k+1]
    ids = np.arange(ns * mm)
    k, c = ids // ns, ids % ns
    assert np.array_equal(d["spatial_ids"], C[c])
    assert np.array_equal(d["temporal_ids"], np.stack([k, k + 1], axis=1))
    # temporal neighbours: the same spatial cell one slab below / above, -1 at t = 0 and T
    tnb = d["temporal_neigh"][:, :, 0]
    assert np.array_equal(tnb[:, 0], np.where(k > 0, ids - ns, -1))
    assert np.array_equal(tnb[:, 1], np.where(k < mm - 1, ids + ns, -1))
    # spatial neighbours: same slab, symmetric, sharing exactly the face's vertices
    snb = d["spatial_neigh"][:, :, 0]
    assert msn == 1
    n_interior = 0
    for cell in range(ns * mm):
        for f in range(nfs):
            other = snb[cell, f]
            if other < 0:
                continue
            n_interior += 1
            assert other // ns == cell // ns
            assert cell in snb[other]
            # face f of a triangle is the edge (v_f, v_(f+1) mod 3) (src/mesh/topology/faces_2d.hpp:15-37),
            # of an interval its vertex f (faces_1d.hpp:15-25)
            tri = C[cell % ns]
            face = {tri[f], tri[(f + 1) % 3]} if dim == 2 else {tri[f]}
            assert face <= set(C[other % ns])
    # each interior spatial facet is listed from both sides, in every slab
    if dim == 2:
        edges = {}
        for tri in C:
            for a, b in ((tri[1], tri[2]), (tri[2], tri[0]), (tri[0], tri[1])):
                key = (min(a, b), max(a, b))
                edges[key] = edges.get(key, 0) + 1
        interior = sum(1 for v in edges.values() if v == 2)
    else:
        interior = ns - 1
    assert n_interior == 2 * interior * mm


def _cellwise(path, payload):
    with open(path, "rb") as f:
        h = _header(f, payload, 1)
        n = int(h[0])
        ids = _block(f, np.int32, n)
        blocks = [_block(f, np.float64, n) for _ in range(1 if payload == 1 else int(h[1]))]
        assert np.fromfile(f, dtype=np.uint8).size == 0
    return h, ids, blocks


def test_cellwise_error_file_order_and_ties(tmp_path):
    values = np.array([0.5, 2.0, 0.5, 1.0, 2.0, 0.0, 0.5])
    path = tmp_path / "cellwise_error.bin"
    snapshot.write_cellwise_error_binary(values, path)
    h, ids, (v,) = _cellwise(path, 1)
    assert int(h[0]) == len(values)
    # value descending, ties to the lower id (time_slab_cellwise_errors.hpp:54-70)
    assert ids.tolist() == [1, 4, 3, 0, 2, 6, 5]
    assert np.array_equal(v, values[ids])
    snapshot.write_cellwise_error_binary(values, path, sort_descending=False)
    h, ids, (v,) = _cellwise(path, 1)
    assert ids.tolist() == list(range(len(values)))
    assert np.array_equal(v, values)


def test_cellwise_error_components_file(tmp_path):
    rng = np.random.default_rng(7)
    n = 50
    est = dict(flux_by_cell=rng.random(n), recon_by_cell=rng.random(n) * 1e-3, resid_by_cell=rng.random(n))
    est["flux_by_cell"][10] = est["flux_by_cell"][20]  # a tie in the total
    est["recon_by_cell"][10] = est["recon_by_cell"][20]
    path = tmp_path / "cellwise_error_components.bin"
    snapshot.write_cellwise_error_components_binary(est, path)
    h, ids, (flux, recon, resid, total) = _cellwise(path, 2)
    assert (int(h[0]), int(h[1])) == (n, 4)
    expect_total = combine_estimator(est["flux_by_cell"], est["recon_by_cell"])
    assert np.array_equal(total, expect_total[ids])
    assert np.array_equal(total, flux + recon)
    assert np.array_equal(flux, est["flux_by_cell"][ids])
    assert np.array_equal(recon, est["recon_by_cell"][ids])
    assert np.array_equal(resid, est["resid_by_cell"][ids])
    assert np.all(np.diff(total) <= 0)
    i10, i20 = list(ids).index(10), list(ids).index(20)
    assert i20 == i10 + 1


# ---- 1+1D adaptive spaces with the reference's cell ids and DoFs ----
def read_dofs(path):
    """as load_dofs_binary (python/binary_readers.py:208-305)"""
    with open(path, "rb") as f:
        nc, dpc, nd, dv, ps, pt, nce = (int(v) for v in _header(f, 7, 1))
        out = dict(n_cells=nc, dofs_per_cell=dpc, n_dofs=nd, dim_v=dv, p_space=ps, p_time=pt,
                   dof_coords=_block(f, np.float64, nd * dv, (nd, dv)),
                   cell_ids=_block(f, np.int32, nc),
                   cell_to_dofs=_block(f, np.int32, nc * dpc, (nc, dpc)),
                   local_reference_coords=_block(f, np.float64, dpc * dv, (dpc, dv)),
                   is_constrained=_block(f, np.uint8, nd).astype(bool),
                   constraint_offsets=_block(f, np.int32, nd + 1),
                   constraint_masters=_block(f, np.int32, nce),
                   constraint_weights=_block(f, np.float64, nce))
        assert np.fromfile(f, dtype=np.uint8).size == 0
    return out


def _adaptive_space(p, seed=3):
    from doubleadapparabolicfem_amd import adaptive1d as A

    rng = np.random.default_rng(seed)
    m = A.Mesh1D()
    X = A.FESpace1D(m, p, True, m.leaf_cells())
    X.refine(X.active)
    for _ in range(4):
        X.refine(rng.choice(X.active, size=max(1, X.n_cells // 4), replace=False))
    return m, X


@pytest.mark.parametrize("p", [1, 2])
def test_adaptive_dofs_file(tmp_path, p):
    m, X = _adaptive_space(p)
    assert X.n_constrained > 0
    path = tmp_path / "dofs.bin"
    snapshot.write_dofs_binary(X, path)
    d = read_dofs(path)
    assert (d["n_cells"], d["dofs_per_cell"], d["dim_v"], d["p_space"], d["p_time"]) == (X.n_cells, (p + 1) ** 2, 2, p, p)
    assert d["n_dofs"] == X.n_dofs + X.n_constrained
    np.testing.assert_array_equal(d["cell_ids"], X.active)
    np.testing.assert_array_equal(d["cell_to_dofs"], X.cell_dofs)
    # every cell's local DoFs sit at the coordinates the file gives their global DoF
    for a in range(X.n_cells):
        x0, x1, t0, t1 = X.box[a]
        for l, (rx, rt) in enumerate(d["local_reference_coords"]):
            g = d["cell_to_dofs"][a, l]
            if g >= 0:
                assert tuple(d["dof_coords"][g]) == (x0 + (x1 - x0) * rx, t0 + (t1 - t0) * rt)
    # constrained DoFs: flag and masters / weights reproduce the restriction rows
    assert not d["is_constrained"][:X.n_dofs].any() and d["is_constrained"][X.n_dofs:].all()
    nl = (p + 1) ** 2
    for a in range(X.n_cells):
        for l in range(nl):
            g = d["cell_to_dofs"][a, l]
            lo, hi = X.restr_ptr[a * nl + l], X.restr_ptr[a * nl + l + 1]
            if g < 0:
                assert hi == lo
            elif not d["is_constrained"][g]:
                assert list(X.restr_dof[lo:hi]) == [g] and list(X.restr_w[lo:hi]) == [1.0]
            else:
                o0, o1 = d["constraint_offsets"][g], d["constraint_offsets"][g + 1]
                assert list(d["constraint_masters"][o0:o1]) == list(X.restr_dof[lo:hi])
                assert list(d["constraint_weights"][o0:o1]) == list(X.restr_w[lo:hi])


def test_adaptive_mesh_file(tmp_path):
    m, X = _adaptive_space(1)
    path = tmp_path / "mesh.bin"
    snapshot.write_space_mesh_binary(X, path)
    d = read_mesh(path)
    h = d["header"]
    assert list(h[2:11]) == [1, 1, X.n_cells, 1, 2, 2, 2, 2, int(h[10])] and h[12] == 2
    x, t = m.vertices()
    np.testing.assert_array_equal(d["spatial_vertices"][:, 0], x)
    np.testing.assert_array_equal(d["temporal_vertices"][:, 0], t)
    np.testing.assert_array_equal(d["cell_ids"], X.active)
    row = {int(c): r for r, c in enumerate(d["cell_ids"])}
    sn, tn = d["spatial_neigh"], d["temporal_neigh"]
    assert sn.shape[2] >= 2  # a hanging face lists both fine neighbours
    for r, c in enumerate(d["cell_ids"]):
        x0, x1 = x[d["spatial_ids"][r]]
        t0, t1 = t[d["temporal_ids"][r]]
        for f, xf in ((0, x0), (1, x1)):
            for nbr in sn[r, f]:
                if nbr < 0:
                    continue
                q = row[int(nbr)]
                qx = x[d["spatial_ids"][q]]
                qt = t[d["temporal_ids"][q]]
                assert xf in qx and max(t0, qt[0]) < min(t1, qt[1])  # shares the face, overlaps in time
                assert c in sn[q, 1 - f]
        for f, tf in ((0, t0), (1, t1)):
            for nbr in tn[r, f]:
                if nbr < 0:
                    continue
                q = row[int(nbr)]
                assert tf in t[d["temporal_ids"][q]]
                assert c in tn[q, 1 - f]


def test_adaptive_function_files(tmp_path):
    m, X = _adaptive_space(2)
    rng = np.random.default_rng(0)
    coef = rng.standard_normal(X.n_dofs)
    snapshot.write_function_binary(X, coef, tmp_path / "f.bin", samples_x=5, samples_t=4)
    snapshot.write_function_true_coefficients_binary(coef, tmp_path / "c.bin")
    with open(tmp_path / "f.bin", "rb") as f:
        nc, ds, dt, nx, nt, vpc = (int(v) for v in _header(f, 6, 1))
        assert (nc, ds, dt, nx, nt, vpc) == (X.n_cells, 1, 1, 5, 4, 20)
        ids = _block(f, np.int32, nc)
        vals = _block(f, np.float64, nc * vpc, (nc, nt, nx))
        assert np.fromfile(f, dtype=np.uint8).size == 0
    np.testing.assert_array_equal(ids, X.active)
    loc = X.local_coefficients(coef)
    # at the sample (xi, eta) = (0, 0) and (1, 1) the value is the local DoF there
    np.testing.assert_allclose(vals[:, 0, 0], loc[:, 0], rtol=0, atol=1e-13)
    np.testing.assert_allclose(vals[:, -1, -1], loc[:, -1], rtol=0, atol=1e-13)
    with open(tmp_path / "c.bin", "rb") as f:
        n, = _header(f, 1, 1)
        np.testing.assert_array_equal(_block(f, np.float64, int(n)), coef)
        assert np.fromfile(f, dtype=np.uint8).size == 0


def _adaptive_pair_1d(p, steps):
    """library X / Y spaces after `steps` outer iterations of the published marking (oracle
    solve), with the oracle's Y space beside them"""
    from doubleadapparabolicfem_amd import adaptive1d as A
    from oracle import adaptive1d as O
    u0 = lambda x: np.sin(np.pi * x)
    T = O.Tree()
    X = O.Space(T, p, True, T.leaves())
    X.refine(list(X.active))
    m = A.Mesh1D()
    LX = A.FESpace1D(m, p, True, m.leaf_cells())
    LX.refine(LX.active)
    for _ in range(steps):
        Y = O.Space(T, p, False, X.active)
        K, b = O.assemble(X, Y, u0)
        x = O.solve(K, b)
        l2, t2 = O.x_indicators(X, Y, x, u0)
        lx = np.zeros(len(X.active))
        np.add.at(lx, Y.ancestors_in(X), l2)
        mk, _ = O.doerfler_mark(lx + t2, 0.4, ids=X.active)
        X.refine(mk)
        LX.refine(mk)
    Y = O.Space(T, p, False, X.active)
    K, b = O.assemble(X, Y, u0)
    return LX, A.FESpace1D(m, p, False, LX.active), Y, O.solve(K, b)


@pytest.mark.parametrize("p", [1, 2])
def test_time_slab_space_and_files_1d(tmp_path, p):
    """TimeSlabSpace of an adaptive Y (time_slab_builder.hpp:53-200): the slab times are the
    cells' distinct time endpoints, slab k holds the active cells spanning it in active order
    (the same cells, left to right, as the oracle's slab partition), each slab space numbers its
    DoFs by first occurrence with the spatial boundary eliminated, (n_cells p - 1)(p + 1) true
    DoFs; the restriction of a Y function is exact; the files carry the reference's layout."""
    from doubleadapparabolicfem_amd import adaptive1d as A
    from oracle import adaptive1d as O
    LX, LY, OY, x = _adaptive_pair_1d(p, 8 if p == 1 else 6)
    tss = A.TimeSlabSpace(LY)
    og = O.slab_groups(OY)
    oslabs = [src for g in og for src in g["src"]]
    assert tss.n_slabs == len(oslabs) and tss.n_slabs > 4
    pos = {int(c): i for i, c in enumerate(LY.active)}
    for sl, osrc in zip(tss.slabs, oslabs):
        ps = [pos[int(c)] for c in sl["source_cells"]]
        assert ps == sorted(ps)
        assert sorted(ps) == sorted(osrc)
        sp_ = sl["space"]
        n = len(ps)
        assert sp_.n_dofs == (n * p - 1) * (p + 1) and sp_.n_constrained == 0
        # first-occurrence numbering by node (x, time index); nodes on x = 0, 1 eliminated
        seen, nxt = {}, 0
        for i in range(n):
            x0, x1 = sp_.box[i, :2]
            for l in range(sp_.n_loc):
                a, j = divmod(l, p + 1)
                xx = x0 + (x1 - x0) * a / p
                if xx in (0.0, 1.0):
                    assert sp_.cell_dofs[i, l] == -1
                    continue
                key = (round(xx * 2**40), j)
                if key not in seen:
                    seen[key] = nxt
                    nxt += 1
                assert sp_.cell_dofs[i, l] == seen[key]
    coefs = tss.restrict(LY, x[:LY.n_dofs])
    lam_loc = LY.local_coefficients(x[:LY.n_dofs])
    L = A._lagrange_values
    rng = np.random.default_rng(1)
    for k, sl in enumerate(tss.slabs):
        loc = sl["space"].local_coefficients(coefs[k])
        for i, c in enumerate(sl["source_cells"]):
            a = pos[int(c)]
            s_, t_ = rng.random(), rng.random()
            bx, bs = sl["space"].box[i], LY.box[a]
            xx, tt = bx[0] + (bx[1] - bx[0]) * s_, bx[2] + (bx[3] - bx[2]) * t_
            v1 = (L(p, [s_]) @ loc[i].reshape(p + 1, p + 1) @ L(p, [t_]).T).item()
            v2 = (L(p, [(xx - bs[0]) / (bs[1] - bs[0])]) @ lam_loc[a].reshape(p + 1, p + 1)
                  @ L(p, [(tt - bs[2]) / (bs[3] - bs[2])]).T).item()
            assert abs(v1 - v2) <= 1e-12 * max(1.0, abs(v2))
    snapshot.write_time_slab_function_binary(tss, coefs, str(tmp_path))
    meta = np.fromfile(tmp_path / "time_slab_space.bin", dtype=np.int32, count=3)
    assert list(meta) == [1, tss.n_slabs, tss.n_slabs + 1]
    times = np.fromfile(tmp_path / "time_slab_space.bin", dtype=np.float64, offset=12)
    np.testing.assert_array_equal(times, tss.times)
    for k in (0, tss.n_slabs - 1):
        d = tmp_path / ("slab_%04d" % k)
        sl = tss.slabs[k]
        n = len(sl["source_cells"])
        raw = (d / "provenance.bin").read_bytes()
        head = np.frombuffer(raw[:12], dtype=np.int32)
        assert list(head) == [1, k, n]
        tb = np.frombuffer(raw[12:28], dtype=np.float64)
        assert list(tb) == [sl["t_begin"], sl["t_end"]]
        ids = np.frombuffer(raw[28:28 + 4 * n], dtype=np.int32)
        src = np.frombuffer(raw[28 + 4 * n:28 + 8 * n], dtype=np.int32)
        iv = np.frombuffer(raw[28 + 8 * n:], dtype=np.float64)
        assert list(ids) == list(range(n)) and list(src) == list(sl["source_cells"])
        assert len(iv) == 2 * n and np.all(iv[0::2] == sl["t_begin"]) and np.all(iv[1::2] == sl["t_end"])
        for name in ("mesh.bin", "dofs.bin", "function.bin", "true_coefficients.bin"):
            assert (d / name).stat().st_size > 0
        tc = np.fromfile(d / "true_coefficients.bin", dtype=np.int32, count=2)
        assert list(tc) == [1, len(coefs[k])]


def test_prolong_nodal_onto_finer_cells_matches_oracle():
    """adaptive1d.prolong_nodal from X onto a degree-(p+1) space on finer cells (Y refined
    past X, the G space of an inner iteration that refined Y) equals the oracle's
    prolongation (fespace/prolongation.hpp:15-97) to round-off."""
    from doubleadapparabolicfem_amd import adaptive1d as A
    from oracle import adaptive1d as O
    p = 1
    LX, LY, OY, x = _adaptive_pair_1d(p, 5)
    T = OY.tree
    OX = O.Space(T, p, True, OY.active)
    marks = [int(c) for c in LY.active[::3]]
    LY.refine(marks)
    OY.refine(marks)
    LG = A.FESpace1D(LY.mesh, p + 1, False, LY.active)
    OG = O.Space(T, p + 1, False, OY.active)
    u = x[-LX.n_dofs:]
    np.testing.assert_allclose(A.prolong_nodal(LX, u, LG), O.prolong_nodal(OX, u, OG), rtol=1e-14, atol=1e-15)


def test_time_slab_edge_patches_1d(tmp_path):
    """TimeSlabSpace.edge_patches (time_slab_edge_patch_builder.hpp:349-527) and its file: per
    slab n_cells + 1 patches (two boundary patches, one interior patch per shared vertex),
    consecutive ids, every slab cell in one or two patches, the interior patch's vertex the
    left cell's right vertex; the file parses back to the same arrays in the reference's
    block order."""
    from doubleadapparabolicfem_amd import adaptive1d as A
    LX, LY, OY, x = _adaptive_pair_1d(2, 6)
    tss = A.TimeSlabSpace(LY)
    P = tss.edge_patches()
    assert [q["patch_id"] for q in P] == list(range(len(P)))
    for k, sl in enumerate(tss.slabs):
        mine = [q for q in P if q["slab_id"] == k]
        n = len(sl["source_cells"])
        assert len(mine) == n + 1
        assert [q["kind"] for q in mine] == [1] + [0] * (n - 1) + [1]
        assert mine[0]["x_vertex"] == 0.0 and mine[-1]["x_vertex"] == 1.0
        counts = np.zeros(n, int)
        for q in mine:
            for c in q["cells"]:
                counts[c["slab_cell_id"]] += 1
            if q["kind"] == 0:
                a, b = q["cells"]
                assert a["right_vertex_id"] == b["left_vertex_id"] == q["spatial_vertex_id"]
                assert (a["side"], b["side"]) == (0, 1) and a["x_end"] == b["x_begin"] == q["x_vertex"]
        assert np.all(counts == 2)
        xs = [q["x_vertex"] for q in mine]
        assert xs == sorted(xs)
    path = tmp_path / "time_slab_edge_patches.bin"
    snapshot.write_time_slab_edge_patches_binary(tss, str(path))
    raw = path.read_bytes()
    o = 0

    def take(dt, count):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt, count=count, offset=o)
        o += a.nbytes
        return a

    head = take(np.int32, 4)
    ns, npch = tss.n_slabs, len(P)
    assert list(head) == [1, ns, npch, 2]
    offs = take(np.int32, ns + 1)
    assert offs[-1] == npch and list(take(np.int32, npch)) == list(range(npch))
    assert list(take(np.int32, npch)) == list(range(npch))
    assert list(take(np.int32, npch)) == [q["slab_id"] for q in P]
    assert list(take(np.int32, npch)) == [q["spatial_vertex_id"] for q in P]
    assert list(take(np.int32, npch)) == [q["kind"] for q in P]
    assert list(take(np.int32, npch)) == [len(q["cells"]) for q in P]
    assert list(take(np.float64, npch)) == [q["x_vertex"] for q in P]
    take(np.float64, npch)
    take(np.float64, npch)
    slab_cell = take(np.int32, 2 * npch)
    assert slab_cell[1] == -1 and slab_cell[0] == P[0]["cells"][0]["slab_cell_id"]
    for _ in range(4):
        take(np.int32, 2 * npch)
    take(np.float64, 2 * npch)
    take(np.float64, 2 * npch)
    assert o == len(raw)


def test_time_slab_vertex_patches_2d(tmp_path):
    """write_time_slab_vertex_patches_binary on a tensor 2+1D mesh: one patch per (slab, vertex),
    the vertex's triangles with the vertex at the recorded local index, boundary vertices of the
    unit square marked, interior patches of the union-jack mesh holding 4 or 8 triangles; the file
    parses back in the reference's block order."""
    m = meshes.uniform_mesh(2, 1)
    dim, V, C, T = snapshot._mesh_arrays(m)
    path = tmp_path / "time_slab_vertex_patches.bin"
    snapshot.write_time_slab_vertex_patches_binary(m, str(path))
    raw = path.read_bytes()
    o = 0

    def take(dt, count):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt, count=count, offset=o)
        o += a.nbytes
        return a

    head = take(np.int32, 4)
    ns_, nslab, nv = len(C), len(T) - 1, len(V)
    assert list(head[:3]) == [2, nslab, nslab * nv]
    n, mc = int(head[2]), int(head[3])
    offs = take(np.int32, nslab + 1)
    assert list(offs) == [k * nv for k in range(nslab + 1)]
    take(np.int32, n)
    assert list(take(np.int32, n)) == list(range(n))
    slab = take(np.int32, n)
    vid = take(np.int32, n)
    kind = take(np.int32, n)
    ncell = take(np.int32, n)
    xv = take(np.float64, 2 * n).reshape(-1, 2)
    tb, te = take(np.float64, n), take(np.float64, n)
    sc, src, lv, lv2, sv = (take(np.int32, n * mc) for _ in range(5))
    tri, tri2 = take(np.int32, 3 * n * mc), take(np.int32, 3 * n * mc)
    xy, ti = take(np.float64, 6 * n * mc), take(np.float64, 2 * n * mc)
    assert o == len(raw)
    on_bnd = (np.abs(V[:, 0]) < 1e-14) | (np.abs(V[:, 0] - 1) < 1e-14) | (np.abs(V[:, 1]) < 1e-14) | \
        (np.abs(V[:, 1] - 1) < 1e-14)
    assert np.array_equal(kind.astype(bool), on_bnd[vid])
    assert set(ncell[kind == 0]) <= {4, 8}
    np.testing.assert_array_equal(xv, V[vid])
    np.testing.assert_array_equal(tb, T[slab]) and np.testing.assert_array_equal(te, T[slab + 1])
    for q in range(n):
        for j in range(mc):
            e = q * mc + j
            if j >= ncell[q]:
                assert sc[e] == -1
                continue
            c = sc[e]
            assert src[e] == slab[q] * ns_ + c and C[c, lv[e]] == vid[q] and sv[e] == vid[q]
            assert list(tri[3 * e:3 * e + 3]) == list(C[c]) and list(tri2[3 * e:3 * e + 3]) == list(C[c])
            np.testing.assert_array_equal(xy[6 * e:6 * e + 6], V[C[c]].reshape(-1))
            assert list(ti[2 * e:2 * e + 2]) == [T[slab[q]], T[slab[q] + 1]]
