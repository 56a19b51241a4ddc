"""Binary snapshots in the reference's file formats.

The writers below produce the files that the reference's readers and plotters
load: the mesh file of ``load_mesh_binary`` (python/binary_readers.py:90-205,
written by src/finite_element/fespace/io/write_mesh_binary.hpp:16-87), and the
per-cell estimator files (src/finite_element/io/write_cellwise_error_binary.hpp:25-78
and write_cellwise_error_components_binary.hpp:14-87).  Every file starts with an
int32 header whose first entry is the format version
(src/finite_element/io/detail/binary_format_versions.hpp:10-22,
``make_versioned_header`` at :28-40); the blocks that follow are raw
little-endian arrays, int32 for ids and float64 for coordinates and values.

Cell ids: the space-time cell of slab k and spatial cell c has id k * ns + c,
the order of this package's per-cell estimator arrays.  (The reference numbers
the active cells of its refinement forest; on the tensor meshes here these ids
take that role.)  Spatial face f of a triangle is the edge (v_f, v_{f+1 mod 3}), as
fill_spatial_faces_2d sets it (src/mesh/topology/faces_2d.hpp:15-37), and face f
of an interval is its vertex f (faces_1d.hpp:15-25); temporal face 0 is a cell's bottom (t_k) and face 1
its top.  A face without a neighbour lists -1, as the reference's
``flatten_neighbours`` does (write_mesh_binary_detail.hpp:149-186).
"""
import numpy as np

MESH_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:10
CELLWISE_ERROR_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:20
CELLWISE_ERROR_COMPONENTS_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:21

_TRI_FACE = ((0, 1), (1, 2), (2, 0))  # face f of a triangle: edge (v_f, v_f+1) (faces_2d.hpp:15-37)


def _write(path, blocks):
    with open(path, "wb") as f:
        for b in blocks:
            np.ascontiguousarray(b).tofile(f)


def _mesh_arrays(mesh):
    dim = int(mesh.dim)
    V = np.asarray(mesh.vertices, dtype=np.float64).reshape(-1, dim)
    C = np.asarray(mesh.cells, dtype=np.int32).reshape(-1, dim + 1)
    T = np.asarray(mesh.times if mesh.times is not None else [0.0, 1.0], dtype=np.float64)
    return dim, V, C, T


def cell_ids(mesh):
    """Ids of the space-time cells in this package's order (slab-major)."""
    dim, V, C, T = _mesh_arrays(mesh)
    return np.arange(len(C) * (len(T) - 1), dtype=np.int32)


def _spatial_face_neighbours(dim, C):
    """[ns][faces] index of the spatial cell across each face, -1 on the boundary."""
    ns, nfc = len(C), dim + 1
    key_of = {}
    for c in range(ns):
        for f in range(nfc):
            key = (int(C[c, f]),) if dim == 1 else tuple(sorted((int(C[c, _TRI_FACE[f][0]]), int(C[c, _TRI_FACE[f][1]]))))
            key_of.setdefault(key, []).append((c, f))
    nb = np.full((ns, nfc), -1, dtype=np.int64)
    for owners in key_of.values():
        if len(owners) > 2:
            raise ValueError("a face shared by more than two cells")
        if len(owners) == 2:
            (c1, f1), (c2, f2) = owners
            nb[c1, f1] = c2
            nb[c2, f2] = c1
    return nb


def write_mesh_binary(mesh, path):
    """Mesh file (format version 1) of a tensor space-time mesh."""
    dim, V, C, T = _mesh_arrays(mesh)
    ns, m = len(C), len(T) - 1
    ncell = ns * m
    tp, ip, nfs, nft = dim + 1, 2, dim + 1, 2
    nb = _spatial_face_neighbours(dim, C)
    max_sn = 1 if (nb >= 0).any() else 0
    max_tn = 1 if m > 1 else 0
    k = np.repeat(np.arange(m), ns)
    c = np.tile(np.arange(ns), m)
    spatial_vids = C[c]
    temporal_vids = np.stack([k, k + 1], axis=1)
    sn = np.where(nb[c] >= 0, k[:, None] * ns + nb[c], -1).reshape(ncell, nfs, 1)[:, :, :max_sn]
    below = np.where(k > 0, (k - 1) * ns + c, -1)
    above = np.where(k < m - 1, (k + 1) * ns + c, -1)
    tn = np.stack([below, above], axis=1).reshape(ncell, nft, 1)[:, :, :max_tn]
    header = [MESH_BINARY_FORMAT_VERSION, len(V), m + 1, dim, 1, ncell, dim, tp, tp, ip, nfs, max_sn, nft, max_tn]
    _write(path, [np.asarray(header, dtype=np.int32), V, T.reshape(-1, 1), np.arange(ncell, dtype=np.int32),
                  spatial_vids.astype(np.int32), temporal_vids.astype(np.int32), sn.astype(np.int32),
                  tn.astype(np.int32)])


def _order(ids, total, sort_descending):
    """Entry order: value descending, ties to the lower id
    (CellwiseSquaredError::sorted_descending, time_slab_cellwise_errors.hpp:54-70);
    otherwise ascending id (the reference iterates an unordered map there)."""
    if sort_descending:
        return np.lexsort((ids, -total))
    return np.argsort(ids, kind="stable")


def write_cellwise_error_binary(values, path, sort_descending=True, ids=None):
    """Per-cell squared errors (format version 1): header [version, n], ids, values."""
    values = np.asarray(values, dtype=np.float64).reshape(-1)
    ids = np.arange(len(values), dtype=np.int64) if ids is None else np.asarray(ids, dtype=np.int64)
    o = _order(ids, values, sort_descending)
    _write(path, [np.asarray([CELLWISE_ERROR_BINARY_FORMAT_VERSION, len(values)], dtype=np.int32),
                  ids[o].astype(np.int32), values[o]])


def write_cellwise_error_components_binary(estimate, path, sort_descending=True):
    """Per-cell estimator components (format version 1) of an ``System.estimate``
    result with per-cell arrays: header [version, n, 4], ids, flux^2,
    reconstruction^2, divergence residual^2 and the total flux^2 + reconstruction^2
    (the combination of apf_combine_estimator), ordered by the total."""
    from . import combine_estimator

    flux = np.asarray(estimate["flux_by_cell"], dtype=np.float64)
    recon = np.asarray(estimate["recon_by_cell"], dtype=np.float64)
    resid = np.asarray(estimate["resid_by_cell"], dtype=np.float64)
    total = combine_estimator(flux, recon, resid)
    ids = np.arange(len(total), dtype=np.int64)
    o = _order(ids, total, sort_descending)
    _write(path, [np.asarray([CELLWISE_ERROR_COMPONENTS_BINARY_FORMAT_VERSION, len(total), 4], dtype=np.int32),
                  ids[o].astype(np.int32), flux[o], recon[o], resid[o], total[o]])


# ---- 1+1D adaptive spaces (adaptive1d.FESpace1D): the reference's own cell ids and DoFs ----
DOFS_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:11
FUNCTION_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:12
FUNCTION_TRUE_COEFFICIENTS_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:14
TIME_SLAB_SPACE_METADATA_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:15
TIME_SLAB_PROVENANCE_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:16
TIME_SLAB_EDGE_PATCHES_BINARY_FORMAT_VERSION = 1  # binary_format_versions.hpp:17
TIME_SLAB_VERTEX_PATCHES_BINARY_FORMAT_VERSION = 2  # binary_format_versions.hpp:18


def write_space_mesh_binary(space, path):
    """Mesh file (format version 1) of an FE space's active cells, as
    FESpace::write_mesh_binary writes it (fespace/io/write_mesh_binary.hpp:16-87): all mesh
    vertices, the active cells' ids and vertex ids, and per face the neighbours across it
    (several across a hanging face) in the adjacency's order, -1 padded."""
    x, t = space.mesh.vertices()
    sn = space.neighbours(temporal=False)
    tn = space.neighbours(temporal=True)
    cl = space.mesh.cells()
    vids = cl["vertex_ids"][space.active]
    header = [MESH_BINARY_FORMAT_VERSION, len(x), len(t), 1, 1, space.n_cells, 1, 2, 2, 2, 2, sn.shape[2], 2,
              tn.shape[2]]
    _write(path, [np.asarray(header, dtype=np.int32), x.astype(np.float64), t.astype(np.float64),
                  space.active.astype(np.int32), vids[:, :2].astype(np.int32), vids[:, 2:].astype(np.int32),
                  sn.astype(np.int32), tn.astype(np.int32)])


def _local_reference_coords(p):
    """(x, t) of local DoF i_space * (p+1) + j_time (tables/element_tables.hpp:182-236)"""
    return np.array([(i / p, j / p) for i in range(p + 1) for j in range(p + 1)], dtype=np.float64)


def write_dofs_binary(space, path):
    """DoF file (format version 1) of FESpace::write_dofs_binary
    (fespace/io/write_dofs_binary.hpp:12-72, detail/write_dofs_binary_detail.hpp:11-148):
    header [version, n_cells, dofs_per_cell, n_dofs (true + constrained), dim_v = 2, p, p,
    constraint entries], DoF coordinates, cell ids, cell-to-DoF table, local reference
    coordinates, constrained flags (uint8), constraint offsets, masters, weights."""
    p, nl = space.p, space.n_loc
    n_dofs = space.n_dofs + space.n_constrained
    ref = _local_reference_coords(p)
    coords = np.zeros((n_dofs, 2))
    seen = np.zeros(n_dofs, dtype=bool)
    for a in range(space.n_cells):
        x0, x1, t0, t1 = space.box[a]
        for l in range(nl):
            g = space.cell_dofs[a, l]
            if g >= 0 and not seen[g]:
                coords[g] = (x0 + (x1 - x0) * ref[l, 0], t0 + (t1 - t0) * ref[l, 1])
                seen[g] = True
    if not seen.all():
        raise RuntimeError("write_dofs_binary: failed to determine coordinates for a global DoF.")
    is_con = np.zeros(n_dofs, dtype=np.uint8)
    is_con[space.n_dofs:] = 1
    offsets = np.zeros(n_dofs + 1, dtype=np.int32)
    offsets[space.n_dofs + 1:] = space.con_ptr[1:]
    header = [DOFS_BINARY_FORMAT_VERSION, space.n_cells, nl, n_dofs, 2, p, p, len(space.con_masters)]
    _write(path, [np.asarray(header, dtype=np.int32), coords, space.active.astype(np.int32),
                  space.cell_dofs.astype(np.int32), ref, is_con, offsets, space.con_masters.astype(np.int32),
                  space.con_weights.astype(np.float64)])


def write_function_binary(space, coef, path, samples_x=32, samples_t=32):
    """Function samples (format version 1, 1+1D) of write_function_binary
    (fespace/io/write_function_binary.hpp:42-92): per active cell the values on a
    samples_t x samples_x reference grid (time outer, space inner)."""
    if samples_x < 2 or samples_t < 2:
        raise ValueError("write_function_binary: need at least 2 samples per direction.")
    p = space.p
    loc = space.local_coefficients(coef)
    xs = np.arange(samples_x) / (samples_x - 1)
    ts = np.arange(samples_t) / (samples_t - 1)
    nodes = np.arange(p + 1) / p

    def lag(s):
        out = np.ones((len(s), p + 1))
        for i in range(p + 1):
            for j in range(p + 1):
                if j != i:
                    out[:, i] *= (s - nodes[j]) / (nodes[i] - nodes[j])
        return out

    Lx, Lt = lag(xs), lag(ts)
    # value(jt, ix) = sum_{i, j} c[i * (p+1) + j] Lx[ix, i] Lt[jt, j]
    vals = np.einsum("cij,xi,tj->ctx", loc.reshape(-1, p + 1, p + 1), Lx, Lt)
    header = [FUNCTION_BINARY_FORMAT_VERSION, space.n_cells, 1, 1, samples_x, samples_t, samples_x * samples_t]
    _write(path, [np.asarray(header, dtype=np.int32), space.active.astype(np.int32), vals.reshape(-1)])


def write_function_true_coefficients_binary(coef, path):
    """True coefficients (format version 1): header [version, n], values
    (write_function_true_coefficients_binary, fespace/io/write_function_binary.hpp:176-196)."""
    coef = np.asarray(coef, dtype=np.float64).reshape(-1)
    _write(path, [np.asarray([FUNCTION_TRUE_COEFFICIENTS_BINARY_FORMAT_VERSION, len(coef)], dtype=np.int32), coef])


def _time_slab_dir(out_dir, k):
    """slab_%04d (io/detail/time_slab_io_detail.hpp:9-22)"""
    import os
    d = os.path.join(out_dir, "slab_%04d" % k)
    os.makedirs(d, exist_ok=True)
    return d


def write_time_slab_space_binary(tss, out_dir, metadata_filename="time_slab_space.bin", mesh_filename="mesh.bin",
                                 dofs_filename="dofs.bin", provenance_filename="provenance.bin"):
    """The files of write_time_slab_space_binary (io/write_time_slab_space_binary.hpp:14-118)
    for an adaptive1d.TimeSlabSpace: the metadata file [version, n_slabs, n_slab_times] and
    the slab times; per slab, in slab_%04d/, the slab space's mesh and DoF files and the
    provenance file [version, slab_id, n_cells], [t_begin, t_end], the slab-local cell ids,
    the source cell ids and each slice's time interval."""
    import os
    os.makedirs(out_dir, exist_ok=True)
    _write(os.path.join(out_dir, metadata_filename),
           [np.asarray([TIME_SLAB_SPACE_METADATA_BINARY_FORMAT_VERSION, tss.n_slabs, len(tss.times)], dtype=np.int32),
            np.asarray(tss.times, dtype=np.float64)])
    for k, sl in enumerate(tss.slabs):
        d = _time_slab_dir(out_dir, k)
        write_space_mesh_binary(sl["space"], os.path.join(d, mesh_filename))
        write_dofs_binary(sl["space"], os.path.join(d, dofs_filename))
        n = len(sl["source_cells"])
        _write(os.path.join(d, provenance_filename),
               [np.asarray([TIME_SLAB_PROVENANCE_BINARY_FORMAT_VERSION, k, n], dtype=np.int32),
                np.asarray([sl["t_begin"], sl["t_end"]], dtype=np.float64),
                np.arange(n, dtype=np.int32), np.asarray(sl["source_cells"], dtype=np.int32),
                np.tile(np.asarray([sl["t_begin"], sl["t_end"]], dtype=np.float64), n)])


def write_time_slab_function_binary(tss, slab_coefs, out_dir, samples_x=32, samples_t=32, **names):
    """The files of write_time_slab_function_binary (io/write_time_slab_function_binary.hpp:10-53):
    the time-slab space files, then per slab the function samples and true coefficients of
    that slab's function (slab_coefs[k], e.g. TimeSlabSpace.restrict of a Y function)."""
    import os
    write_time_slab_space_binary(tss, out_dir, **{k: v for k, v in names.items() if k.endswith("_filename")
                                                  and k not in ("function_filename", "coefficients_filename")})
    for k, sl in enumerate(tss.slabs):
        d = _time_slab_dir(out_dir, k)
        write_function_binary(sl["space"], slab_coefs[k], os.path.join(d, names.get("function_filename", "function.bin")),
                              samples_x, samples_t)
        write_function_true_coefficients_binary(
            slab_coefs[k], os.path.join(d, names.get("coefficients_filename", "true_coefficients.bin")))


def write_time_slab_edge_patches_binary(tss, path):
    """The 1+1D edge-patch file (format version 1, io/write_time_slab_edge_patches_binary.hpp:
    14-155 with detail/time_slab_patch_io_detail.hpp:9-103) of TimeSlabSpace.edge_patches():
    header [version, n_slabs, n_patches, 2]; the slab index (offsets [n_slabs+1], patch ids);
    per patch its id, slab, spatial vertex, kind (0 interior, 1 boundary), cell count, vertex
    x, t_begin, t_end; per patch cell (2 per patch, -1 / 0.0 padded) the slab cell, source
    cell, left and right vertex ids, side (0 left of the vertex, 1 right) and x extent."""
    patches = tss.edge_patches()
    n, mc = len(patches), 2
    offsets = [0]
    ids = []
    for k in range(tss.n_slabs):
        ids += [q["patch_id"] for q in patches if q["slab_id"] == k]
        offsets.append(len(ids))
    ci = {key: np.full(n * mc, -1, dtype=np.int32) for key in ("slab_cell_id", "source_cell_id", "left_vertex_id",
                                                                "right_vertex_id", "side")}
    cx = {key: np.zeros(n * mc) for key in ("x_begin", "x_end")}
    for q in patches:
        for j, c in enumerate(q["cells"]):
            for key in ci:
                ci[key][q["patch_id"] * mc + j] = c[key]
            for key in cx:
                cx[key][q["patch_id"] * mc + j] = c[key]
    col = lambda key, dt: np.asarray([q[key] for q in patches], dtype=dt)
    _write(path, [np.asarray([TIME_SLAB_EDGE_PATCHES_BINARY_FORMAT_VERSION, tss.n_slabs, n, mc], dtype=np.int32),
                  np.asarray(offsets, dtype=np.int32), np.asarray(ids, dtype=np.int32),
                  col("patch_id", np.int32), col("slab_id", np.int32), col("spatial_vertex_id", np.int32),
                  col("kind", np.int32), np.asarray([len(q["cells"]) for q in patches], dtype=np.int32),
                  col("x_vertex", np.float64), col("t_begin", np.float64), col("t_end", np.float64),
                  ci["slab_cell_id"], ci["source_cell_id"], ci["left_vertex_id"], ci["right_vertex_id"], ci["side"],
                  cx["x_begin"], cx["x_end"]])


def write_time_slab_vertex_patches_binary(mesh, path):
    """The 2+1D vertex-patch file (format version 2, io/write_time_slab_vertex_patches_binary.hpp:
    10-216 with detail/time_slab_patch_io_detail.hpp:9-103) of a tensor mesh T_h x {t_k}: slab k
    is the interval [t_k, t_k+1] over the spatial mesh, one patch per (slab, spatial vertex)
    (vertices ascending, the triangles holding the vertex ascending), kind 1 on the boundary.
    Cell ids are this package's (slab cell = spatial triangle c, source cell k * ns + c), as in
    write_mesh_binary.  Blocks: header [version, n_slabs, n_patches, max cells]; slab index;
    per patch id, slab, vertex, kind, cell count, vertex (x, y), t_begin, t_end; per patch cell
    (max cells per patch, -1 / 0.0 padded) slab cell, source cell, local vertex index (twice),
    source vertex, the triangle's vertex ids (twice), its vertex coordinates and time interval."""
    dim, V, C, T = _mesh_arrays(mesh)
    if dim != 2:
        raise ValueError("write_time_slab_vertex_patches_binary: 2+1D meshes only")
    ns, m, nv = len(C), len(T) - 1, len(V)
    cells_of = [[] for _ in range(nv)]
    for c in range(ns):
        for a in range(3):
            cells_of[int(C[c, a])].append(c)
    edge_count = {}
    for c in range(ns):
        for a, b in _TRI_FACE:
            key = tuple(sorted((int(C[c, a]), int(C[c, b]))))
            edge_count[key] = edge_count.get(key, 0) + 1
    boundary = np.zeros(nv, dtype=bool)
    for (a, b), n in edge_count.items():
        if n == 1:
            boundary[a] = boundary[b] = True
    mc = max(len(x) for x in cells_of)
    verts = [v for v in range(nv) if cells_of[v]]
    n = m * len(verts)
    pid = np.arange(n, dtype=np.int32)
    slab = np.repeat(np.arange(m, dtype=np.int32), len(verts))
    vid = np.tile(np.asarray(verts, dtype=np.int32), m)
    kind = boundary[vid].astype(np.int32)
    ncell = np.asarray([len(cells_of[v]) for v in vid], dtype=np.int32)
    xv = V[vid].reshape(-1)
    tb, te = T[slab].astype(np.float64), T[slab + 1].astype(np.float64)
    flat = n * mc
    sc, src, lv, sv = (np.full(flat, -1, dtype=np.int32) for _ in range(4))
    tri = np.full(flat * 3, -1, dtype=np.int32)
    xy = np.zeros(flat * 6)
    ti = np.zeros(flat * 2)
    for q in range(n):
        v, k = int(vid[q]), int(slab[q])
        for j, c in enumerate(sorted(cells_of[v])):
            e = q * mc + j
            sc[e], src[e], sv[e] = c, k * ns + c, v
            lv[e] = int(np.flatnonzero(C[c] == v)[0])
            tri[3 * e:3 * e + 3] = C[c]
            xy[6 * e:6 * e + 6] = V[C[c]].reshape(-1)
            ti[2 * e:2 * e + 2] = (T[k], T[k + 1])
    offsets = np.arange(0, n + 1, len(verts), dtype=np.int32)
    _write(path, [np.asarray([TIME_SLAB_VERTEX_PATCHES_BINARY_FORMAT_VERSION, m, n, mc], dtype=np.int32),
                  offsets, pid, pid, slab, vid, kind, ncell, xv, tb, te, sc, src, lv, lv, sv, tri, tri, xy, ti])
