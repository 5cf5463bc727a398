"""GPU parity of the native path: lattice field and indexed extraction.

The HIP results are compared with the oracle (oracle/oracle.cpp) on the same
inputs:
  * field values: bit-exact expected, at most 4 ulp allowed;
  * sign words, vertex count, triangle count, vertex positions and triangle
    indices: bit-exact (extraction is fed the GPU's own field);
  * full-size workloads: size-independent properties (closed 2-manifold,
    consistent winding, index range) plus oracle slabs of the field.
"""
import numpy as np
import pytest
import torch

from gpucadforam_amd import configs as CF
from tests.helpers import (edge_use_counts, euler_characteristic, mesh_slab_against_oracle,
                           oriented_consistently, to_np, ulp_diff)

pytestmark = pytest.mark.gpu

FIELD_ULP = 4


def _case_list():
    cases = [
        ("plain_gyroid_128", CF.plain_gyroid(128)),
        ("varying_schwarz_p_64", CF.varying_schwarz_p(64)),
        ("varying_schwarz_p_97", CF.varying_schwarz_p(97)),
        ("bent_diamond_96", CF.bent_diamond(96)),
    ]
    for i, p in enumerate(CF.batched_grids(6, seed=0xBA7C)):
        cases.append(("batched_%d_%dx%dx%d" % (i, p.nx, p.ny, p.nz), p))
    for i, p in enumerate(CF.served_requests(6)):
        cases.append(("served_%d_%d" % (i, p.nx), p))
    # period ramp without rotation, gyroid with rotation, diamond with rotation
    cases.append(("ramp_gyroid", CF.make(70, 66, 40, 0, 10.0, 30.0, sheet=True, t0=0.2)))
    cases.append(("rot_gyroid", CF.make(72, 40, 33, 0, 12.0, rot0=0.3, rot1=1.2, sheet=True,
                                        t0=0.25, modes=CF.fourier_modes(7))))
    cases.append(("rot_diamond", CF.make(65, 64, 31, 2, 14.0, 20.0, rot0=-0.5, rot1=0.5,
                                         sheet=False, prim=2, op=3, half=(0.3, 0.35, 0.4))))
    # squared primitive distances of 0 and below 2^-96, where the field kernel
    # takes its rescaled square root: a sphere centred on grid point
    # (32, 32, 16), a box whose inside is at distance 0, and a centre 1e-20
    # from the origin (subnormal squares at the first points)
    cases.append(("sphere_on_point", CF.make(65, 65, 33, 1, 12.0, 20.0, rot0=0.1, rot1=0.9,
                                             sheet=True, t0=0.3, prim=1, op=1,
                                             center=(0.5, 0.5, 0.5), radius=0.3)))
    cases.append(("box_inside", CF.make(65, 65, 33, 2, 14.0, rot0=0.2, rot1=0.6, sheet=True,
                                        t0=0.2, prim=2, op=3, center=(0.5, 0.5, 0.5),
                                        half=(0.2, 0.2, 0.2))))
    cases.append(("sphere_near_origin", CF.make(64, 40, 24, 0, 10.0, sheet=False, prim=1, op=2,
                                                center=(1e-20, 1e-21, 0.0), radius=0.1)))
    return cases


CASES = _case_list()


def gpu_field(gc, p, with_signs=True):
    f = torch.empty(p.npoints(), dtype=torch.float32, device="cuda")
    s = None
    if with_signs:
        s = torch.zeros(gc.sign_words(p.nx, p.ny, p.nz), dtype=torch.int64, device="cuda")
    gc.lattice_field(p, out=f, signs=s, iso=0.0)
    torch.cuda.synchronize()
    return f, s


@pytest.mark.parametrize("name,p", CASES, ids=[c[0] for c in CASES])
def test_field_matches_oracle(gc, orc, dev, name, p):
    f, s = gpu_field(gc, p)
    ref = orc.lattice_field_f32(p)
    got = to_np(f)
    d = ulp_diff(got, ref)
    assert d.max() <= FIELD_ULP, "%s: max ulp %d at %d" % (name, d.max(), d.argmax())
    # the sign words are the ballot of the same values
    sref = orc.signs(got, p.nx, p.ny, p.nz, 0.0)
    assert np.array_equal(to_np(s).view(np.uint64), sref)


@pytest.mark.parametrize("name,p", CASES, ids=[c[0] for c in CASES])
def test_extract_matches_oracle(gc, orc, dev, name, p):
    f, s = gpu_field(gc, p)
    field = to_np(f)
    v_ref, t_ref, _ = orc.extract_indexed(field, p.nx, p.ny, p.nz)
    for signs in (s, None):  # signs from the field kernel, and computed by extraction
        m = gc.extract(f, p.nx, p.ny, p.nz, signs=signs)
        assert (m.n_vertices, m.n_triangles) == (len(v_ref), len(t_ref)), name
        assert np.array_equal(to_np(m.vertices).view(np.uint32), v_ref.view(np.uint32)), name
        assert np.array_equal(to_np(m.triangles).astype(np.uint32), t_ref), name


def _manual_field(nx, ny, nz, fn):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.ascontiguousarray(fn(x, y, z).astype(np.float32).reshape(-1))


EDGE_FIELDS = [
    ("tiny_2", 2, 2, 2, lambda x, y, z: np.where((x + y + z) % 2 == 0, -1.0, 1.0)),
    ("odd_3x5x7", 3, 5, 7, lambda x, y, z: np.sin(x * 1.3 + y * 0.7) + np.cos(z * 0.9) - 0.2),
    ("nx65", 65, 9, 6, lambda x, y, z: np.sin(x * 0.21) * np.cos(y * 0.5 + z * 0.3)),
    ("nx130", 130, 7, 5, lambda x, y, z: np.cos(x * 0.13 + y) - 0.1 * z),
    ("nx64", 64, 8, 8, lambda x, y, z: (x - 31.5) ** 2 + (y - 3.5) ** 2 + (z - 3.5) ** 2 - 9.0),
    ("nx127", 127, 4, 3, lambda x, y, z: np.sin(0.3 * x) + 0.01 * y - 0.02 * z),
    ("checker", 70, 6, 5, lambda x, y, z: np.where((x + y + z) % 2 == 0, -1.0, 1.0)),
    ("iso_ties", 40, 9, 8, lambda x, y, z: np.round(np.sin(0.4 * x + 0.3 * y + 0.2 * z) * 2) / 2),
    ("all_inside", 33, 9, 9, lambda x, y, z: -1.0 - 0.0 * x),
    ("all_outside", 33, 9, 9, lambda x, y, z: 1.0 + 0.0 * x),
]


@pytest.mark.parametrize("name,nx,ny,nz,fn", EDGE_FIELDS, ids=[e[0] for e in EDGE_FIELDS])
def test_extract_edge_cases(gc, orc, dev, name, nx, ny, nz, fn):
    field = _manual_field(nx, ny, nz, fn)
    f = torch.from_numpy(field).to(dev)
    v_ref, t_ref, _ = orc.extract_indexed(field, nx, ny, nz)
    m = gc.extract(f, nx, ny, nz)
    assert (m.n_vertices, m.n_triangles) == (len(v_ref), len(t_ref))
    assert np.array_equal(to_np(m.vertices).view(np.uint32), v_ref.view(np.uint32))
    assert np.array_equal(to_np(m.triangles).astype(np.uint32), t_ref)


def test_extract_overflow_reports_counts(gc, orc, dev):
    import ctypes as C
    p = CF.plain_gyroid(64)
    f, _ = gpu_field(gc, p, with_signs=False)
    v_ref, t_ref, _ = orc.extract_indexed(to_np(f), 64, 64, 64)
    wsb = gc.extract_workspace_bytes(64, 64, 64)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    vcap, tcap = len(v_ref) // 2, len(t_ref) // 2
    verts = torch.full((vcap + 16, 3), 7.0, dtype=torch.float32, device=dev)
    tris = torch.full((tcap + 16, 3), 7, dtype=torch.int32, device=dev)
    a = gc.ExtractArgs()
    a.field, a.signs = f.data_ptr(), None
    a.nx = a.ny = a.nz = 64
    a.iso = 0.0
    a.origin = gc.Float3(0, 0, 0)
    a.spacing = gc.Float3(1 / 63, 1 / 63, 1 / 63)
    a.vertices, a.vert_capacity = verts.data_ptr(), vcap
    a.triangles, a.tri_capacity = tris.data_ptr(), tcap
    a.normals = None
    a.workspace, a.workspace_bytes = ws.data_ptr(), wsb
    a.d_counts = None
    counts = gc.MeshCounts()
    st = gc.lib.gpucad_extract(C.byref(a), C.byref(counts), gc.stream_handle())
    assert st == gc.E_OVERFLOW
    assert (counts.n_vertices, counts.n_triangles) == (len(v_ref), len(t_ref))
    # nothing written past the capacities; what was written is the prefix
    assert torch.all(verts[vcap:] == 7.0) and torch.all(tris[tcap:] == 7)
    assert np.array_equal(to_np(verts[:vcap]).view(np.uint32), v_ref[:vcap].view(np.uint32))


def test_extract_emit_event(gc, orc, dev):
    """gc_extract_args.emit_event: a HIP event recorded on the stream just
    before the emit launch, between an event before the call and one after
    it; the mesh is the one made without the event."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    ev = [C.c_void_p() for _ in range(3)]
    for e in ev:
        assert hip.hipEventCreate(C.byref(e)) == 0
    n = 96
    p = CF.plain_gyroid(n)
    f, s = gpu_field(gc, p)
    ref = gc.extract(f, n, n, n, signs=s)
    wsb = gc.extract_workspace_bytes(n, n, n)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    verts = torch.empty((ref.n_vertices, 3), dtype=torch.float32, device=dev)
    tris = torch.empty((ref.n_triangles, 3), dtype=torch.int32, device=dev)
    a = gc.ExtractArgs()
    a.field, a.signs = f.data_ptr(), s.data_ptr()
    a.nx = a.ny = a.nz = n
    a.iso = 0.0
    a.origin = gc.Float3(0, 0, 0)
    a.spacing = gc.Float3(1 / (n - 1), 1 / (n - 1), 1 / (n - 1))
    a.vertices, a.vert_capacity = verts.data_ptr(), ref.n_vertices
    a.triangles, a.tri_capacity = tris.data_ptr(), ref.n_triangles
    a.normals = None
    a.workspace, a.workspace_bytes = ws.data_ptr(), wsb
    a.d_counts = None
    a.emit_event = ev[1].value
    h = gc.stream_handle()
    assert hip.hipEventRecord(ev[0], h) == 0
    gc.check(gc.lib.gpucad_extract(C.byref(a), None, h))
    assert hip.hipEventRecord(ev[2], h) == 0
    assert hip.hipEventSynchronize(ev[2]) == 0
    before, emit = C.c_float(), C.c_float()
    assert hip.hipEventElapsedTime(C.byref(before), ev[0], ev[1]) == 0
    assert hip.hipEventElapsedTime(C.byref(emit), ev[1], ev[2]) == 0
    assert before.value >= 0.0 and emit.value > 0.0
    assert torch.equal(verts.view(torch.int32), ref.vertices.view(torch.int32))
    assert torch.equal(tris, ref.triangles.to(torch.int32))
    for e in ev:
        hip.hipEventDestroy(e)


def test_extract_rejects_bad_arguments(gc, dev):
    import ctypes as C
    a = gc.ExtractArgs()
    a.field = None
    counts = gc.MeshCounts()
    assert gc.lib.gpucad_extract(C.byref(a), C.byref(counts), None) == gc.E_ARG
    p = CF.plain_gyroid(16)
    p.nx = 1
    f = torch.empty(16 * 16 * 16, dtype=torch.float32, device=dev)
    with pytest.raises(gc.GpucadError):
        gc.lattice_field(p, out=f)
    p = CF.plain_gyroid(16)
    p.rot1 = 0.5
    p.bend_radius = 1.0  # rotation and bend are exclusive
    with pytest.raises(gc.GpucadError):
        gc.lattice_field(p, out=f)


def test_sphere_mesh_topology(gc, orc, dev):
    """A sphere inside the grid meshes to a closed, consistently wound genus-0 surface."""
    n = 80
    f = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    st = gc.lib.gpucad_sphere_with_center(gc.ptr(f), gc.Float3(0.0, 0.0, 0.0), 25.0, 0.0, n, n, n,
                                          1.0, 1.0, 1.0, 0)
    gc.check(st)
    m = gc.extract(f, n, n, n, spacing=(1.0, 1.0, 1.0), origin=(-39.5, -39.5, -39.5))
    t = to_np(m.triangles)
    assert edge_use_counts(t) == {2: 3 * len(t) // 2}
    chi, all_used = euler_characteristic(m.n_vertices, t)
    assert chi == 2 and all_used
    assert oriented_consistently(t)
    # vertices lie on the sphere |p| = 25 up to the interpolation error of a
    # squared implicit over one cell
    r = np.linalg.norm(to_np(m.vertices), axis=1)
    assert np.abs(r - 25.0).max() < 0.05


def test_linear_field_vertices_exact(gc, orc, dev):
    """For f = a.p + b the interpolated crossing is exact up to rounding."""
    nx, ny, nz = 40, 30, 20
    a = np.array([0.37, -0.61, 0.29])
    b = 0.113
    h = np.array([1 / (nx - 1), 1 / (ny - 1), 1 / (nz - 1)])
    field = _manual_field(nx, ny, nz, lambda x, y, z: a[0] * x * h[0] + a[1] * y * h[1] + a[2] * z * h[2] + b)
    m = gc.extract(torch.from_numpy(field).to(dev), nx, ny, nz)
    v = to_np(m.vertices).astype(np.float64)
    resid = v @ a + b
    assert m.n_triangles > 0
    assert np.abs(resid).max() < 2e-6


def _slab_io(fv, m):
    """Host views of planes of a device field and ranges of a device mesh."""
    planes = lambda a, b: to_np(fv[a:b]).reshape(-1)
    verts = lambda i, j: to_np(m.vertices[i:j])
    tris = lambda i, j: to_np(m.triangles[i:j]).astype(np.uint32)
    return planes, verts, tris


def _edges_on_device(t, nv):
    """Edge-use histogram values and the directed-edge uniqueness of a device
    triangle list (T, 3): (max uses of an undirected edge, undirected edges
    used once as (a, b) pairs on the host, whether every directed edge is
    unique = consistent winding)."""
    t = t.long()
    a = torch.cat([t[:, 0], t[:, 1], t[:, 2]])
    b = torch.cat([t[:, 1], t[:, 2], t[:, 0]])
    d = torch.sort(a * nv + b).values
    wound = bool((d[1:] != d[:-1]).all())
    del d
    u, cnt = torch.unique(torch.minimum(a, b) * nv + torch.maximum(a, b), return_counts=True)
    once = u[cnt == 1]
    return int(cnt.max()), torch.stack([once // nv, once % nv], 1), wound


def test_full_size_varying_schwarz_p(gc, orc, dev):
    """512^3 spatially varying Schwarz-P ∩ sphere: oracle slabs of the field,
    oracle slabs of the mesh (bottom, middle, top: triangle ids and order,
    vertex bits) and the mesh invariants of the whole extraction."""
    p = CF.varying_schwarz_p(512)
    f, s = gpu_field(gc, p)
    fv = f.view(512, 512, 512)
    for z0 in (0, 255, 508):
        ref = orc.lattice_field_f32(p, z0, z0 + 4)
        assert ulp_diff(to_np(fv[z0:z0 + 4]).reshape(-1), ref).max() <= FIELD_ULP
    m = gc.extract(f, 512, 512, 512, signs=s)
    t = m.triangles
    assert m.n_triangles > 1_000_000
    assert int(t.min()) >= 0 and int(t.max()) < m.n_vertices
    sp = (1 / 511, 1 / 511, 1 / 511)
    planes, verts, tris = _slab_io(fv, m)
    # the sphere (r = 0.45) spans planes 26..485: its bottom cap, the middle,
    # and a top slab that reaches the last plane (offsets from the totals)
    for z0, z1 in ((16, 40), (240, 272), (470, 511)):
        print("512^3 mesh slab z %d..%d against the oracle" % (z0, z1), flush=True)
        nt, nv = mesh_slab_against_oracle(orc, planes, 512, 512, 512, sp, z0, z1, verts, tris,
                                          (m.n_vertices, m.n_triangles))
        assert nt > 10_000
    tn = to_np(t)
    assert edge_use_counts(tn) == {2: 3 * len(tn) // 2}  # closed (the sphere is interior)
    assert oriented_consistently(tn)
    # every vertex is used and lies on a grid edge inside the unit cube
    v = to_np(m.vertices)
    assert v.min() >= 0.0 and v.max() <= 1.0
    assert np.unique(tn).size == m.n_vertices


def test_full_size_bent_diamond(gc, orc, dev):
    """1024^3 bent diamond ∪ box (4 GiB field): oracle slabs of the field and
    of the mesh (triangle ids and order, vertex bits), and on the device: every
    vertex used, every edge used once or twice, edges used once only on the
    domain's boundary faces, consistent winding."""
    n = 1024
    p = CF.bent_diamond(n)
    f, s = gpu_field(gc, p)
    fv = f.view(n, n, n)
    for z0 in (0, 700):
        ref = orc.lattice_field_f32(p, z0, z0 + 2)
        assert ulp_diff(to_np(fv[z0:z0 + 2]).reshape(-1), ref).max() <= FIELD_ULP
    m = gc.extract(f, n, n, n, signs=s, vert_capacity=300_000_000, tri_capacity=600_000_000)
    assert m.n_triangles > 10_000_000
    t = m.triangles
    assert int(t.min()) >= 0 and int(t.max()) < m.n_vertices
    sp = (1 / (n - 1),) * 3
    planes, verts, tris = _slab_io(fv, m)
    for z0, z1 in ((0, 16), (500, 520), (1010, 1023)):
        print("1024^3 mesh slab z %d..%d against the oracle" % (z0, z1), flush=True)
        nt, nv = mesh_slab_against_oracle(orc, planes, n, n, n, sp, z0, z1, verts, tris,
                                          (m.n_vertices, m.n_triangles))
        assert nt > 0
    del f, s, fv
    torch.cuda.empty_cache()
    print("1024^3 edge use and winding (%d triangles)" % m.n_triangles, flush=True)
    used = torch.zeros(m.n_vertices, dtype=torch.bool, device=dev)
    used[t.reshape(-1).long()] = True
    assert bool(used.all())
    del used
    max_use, once, wound = _edges_on_device(t, m.n_vertices)
    assert max_use == 2 and wound
    # an edge used by one triangle lies in a boundary face of the domain: both
    # ends share a coordinate equal to 0 or to the last grid plane's
    hi = float(np.float32(np.float32(n - 1) * np.float32(sp[0])))
    va, vb = m.vertices[once[:, 0]], m.vertices[once[:, 1]]
    on = lambda c: ((va[:, c] == 0) & (vb[:, c] == 0)) | ((va[:, c] == hi) & (vb[:, c] == hi))
    assert len(once) > 0 and bool((on(0) | on(1) | on(2)).all())
