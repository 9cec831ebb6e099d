"""Host logic of the TomoKernel.conf runner (no GPU): the filter file
format, TomoKernel.conf, the kernel sub-grid and its blocks, pick windows
and the block-by-block kernel file (SI_ker_sta.F90:135-821, 1047-1125)."""
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from fwant_amd import io as fio
from fwant_amd import kernel_run as kr


def test_filter_file_round_trip_and_fixed_fields(tmp_path):
    b = np.array([0.0200833655642112, 0.0401667311284225, 0.0200833655642112])
    a = np.array([1.0, -1.5610180758007182, 0.6413515380575631])
    p = tmp_path / "f.txt"
    kr.write_filter(p, b, a)
    lines = p.read_text().splitlines()
    assert lines[0] == "3" and all(len(ln) == 40 for ln in lines[1:])
    b2, a2 = kr.read_filter(p)
    np.testing.assert_allclose(b2, b, atol=1e-16)
    np.testing.assert_allclose(a2, a, atol=1e-16)
    # F20.16 input: blank is zero, no decimal point means 16 implied decimals
    assert kr._fortran_f(" " * 20) == 0.0
    assert kr._fortran_f("               12345") == pytest.approx(12345e-16)
    assert kr._fortran_f("  1.5000000000000000") == 1.5


def write_conf(tmp_path, **over):
    keys = dict(MAIN_CONF="SeisFD3D.conf", snap_id=1, first_run="T", sub_start="1 2 1", sub_count="-1 8 -1",
                sub_stride="1 1 2", block_from="1 1 1", block_size="5 4 -1", SGT_TIMEDIM_SIZE=36, SGT_M0="1.0e16",
                SGT_ROOT="sgt", SGT_STAX="x", SGT_STAY="y", SGT_STAZ="z", EVENT_ROOT="event", EVENT_OUTD="out",
                maximum_of_pick_per_component=4, number_of_freqband=2)
    keys.update(over)
    text = "\n".join(f"{k} = {v}  # comment" for k, v in keys.items())
    text += "\n<filter_file>\n  band1.txt\n\nnone\n"
    kr.write_filter(tmp_path / "band1.txt", [0.25, 0.5, 0.25], [1.0, -0.2, 0.1])
    p = tmp_path / "TomoKernel.conf"
    p.write_text(text)
    # a main conf with one stress snapshot: interior start 5 5 7, 12 x 10 x 6 points, stride 2, every 2nd step
    (tmp_path / "SeisFD3D.conf").write_text("ni = 40\nnj = 36\nnk = 32\nstept = 0.05\nMEDIA_ROOT = input/media\n"
                                            "number_of_snap = 1\nsnap_001 = 5 5 7 12 10 6 2 2 2 2 10 T\n")
    return p


def test_kernel_conf_and_grid(tmp_path):
    kc = kr.read_kernel_conf(write_conf(tmp_path))
    assert (kc.kid, kc.first_run, kc.knt, kc.M0) == (1, True, 36, 1e16)
    assert kc.bstart == (1, 2, 1) and kc.bcount == (-1, 8, -1) and kc.bstride == (1, 1, 2)
    assert kc.sgt_sta == ("x", "y", "z") and len(kc.bands) == 2 and kc.bands[1] is None
    np.testing.assert_allclose(kc.bands[0][0], [0.25, 0.5, 0.25])
    g = kr.plan_grid(kc)
    assert g.snap["subs"] == (8, 8, 10) and g.snap["subc"] == (12, 10, 6) and g.tcnt == 10
    assert g.kdt == np.float32(2) * np.float32(0.05)
    # (12 - 1)/1 + 1 = 12; (10 - 2)/1 + 1 = 9 capped at 8; (6 - 1)/2 + 1 = 3
    assert g.subnc_s == (1, 2, 1) and g.subnc_c == (12, 8, 3) and g.subnc_t == (1, 1, 2)
    assert g.blksiz == (5, 4, 3) and g.blknum == (3, 2, 1)


def test_blocks_tile_the_kernel_grid(tmp_path):
    kc = kr.read_kernel_conf(write_conf(tmp_path))
    g = kr.plan_grid(kc)
    seen = []
    order = []
    for b, s, c, t in kr.blocks(kc, g):
        order.append(b)
        for k in range(c[2]):
            for j in range(c[1]):
                for i in range(c[0]):
                    seen.append((s[0] + i * t[0], s[1] + j * t[1], s[2] + k * t[2]))
    want = [(1 + i, 2 + j, 1 + 2 * k) for k in range(3) for j in range(8) for i in range(12)]
    assert sorted(seen) == sorted(want) and len(seen) == len(set(seen))
    assert order == [(1, 1, 1), (2, 1, 1), (3, 1, 1), (1, 2, 1), (2, 2, 1), (3, 2, 1)]
    # block_from skips the earlier blocks of each direction
    kc2 = kr.read_kernel_conf(write_conf(tmp_path, block_from="2 2 1"))
    assert [b for b, *_ in kr.blocks(kc2, kr.plan_grid(kc2))] == [(2, 2, 1), (3, 2, 1)]


def test_pick_window():
    knt, kdt = 36, np.float32(0.1)
    T = np.arange(1, knt + 1, dtype=np.float32) * kdt
    assert kr.pick_window(T, 0.85, 2.95, kdt, knt) == (8, 30)  # last T <= t1, first T >= t2
    assert kr.pick_window(T, -1.0, 1e9, kdt, knt) == (1, knt)  # clamped to the record
    assert kr.pick_window(T, 0.1, 3.6, kdt, knt) == (1, knt)
    assert kr.pick_window(T, 3.65, 4.0, kdt, knt) is None  # t1 > Tmax
    assert kr.pick_window(T, 0.0, 0.05, kdt, knt) is None  # t2 < Tmin
    assert kr.pick_window(T, 1.0, 1.0, kdt, knt) == (10, 10)


def test_event_and_pick_lines():
    name, pk, rot = kr.parse_event_header("EVT1 1 0 2 1 1 0  1 2 3 4 5 6 7 8 9")
    assert name == "EVT1" and pk == (1, 0, 2, 1, 1, 0)
    assert rot[1, 0] == 2 and rot[0, 1] == 4  # column-major rot_mat(3,3)
    p = kr.parse_pick("2 P1 1.5 3.25 out/dir one", 2)
    assert (p.band, p.filt, p.t1, p.t2, p.outdir) == (2, "P1", 1.5, 3.25, "out/dir one")
    with pytest.raises(ValueError, match="overflow"):
        kr.parse_pick("3 P1 1 2 o", 2)
    with pytest.raises(ValueError, match="pickfilt"):
        kr.parse_pick("1 P3 1 2 o", 2)


@pytest.mark.parametrize("version", [1, 2])
def test_classic_layout_reads_scipy_files(tmp_path, version):
    p = tmp_path / "a.nc"
    rng = np.random.default_rng(version)
    u = rng.standard_normal((3, 4, 5)).astype(np.float32)
    w = rng.standard_normal((2, 7)).astype(np.float64)
    with netcdf_file(p, "w", version=version) as nc:
        nc.title = "x" * 7
        nc.createDimension("t", None)
        for d, n in (("K", 3), ("J", 4), ("I", 5), ("a", 2), ("b", 7)):
            nc.createDimension(d, n)
        nc.createVariable("rec", "f4", ("t", "I"))[:2] = np.ones((2, 5), np.float32)
        v = nc.createVariable("u", "f4", ("K", "J", "I"))
        v.units = "m"
        v[:] = u
        nc.createVariable("w", "f8", ("a", "b"))[:] = w
    lay = fio.classic_layout(p)
    assert set(lay) == {"u", "w"}
    for name, want in (("u", u), ("w", w)):
        begin, shape, dt = lay[name]
        got = np.fromfile(p, dtype=dt, count=want.size, offset=begin).reshape(shape)
        np.testing.assert_array_equal(got, want)


def test_grid3d_create_and_block_put(tmp_path):
    p = str(tmp_path / "kernel_001_mpi000000_n00000.nc")
    fio.grid3d_create(p, (12, 10, 6), kr.KERNEL_VARS, "Finite frequency kernel")
    rng = np.random.default_rng(3)
    full = {n: np.zeros((6, 10, 12), np.float32) for n in kr.KERNEL_VARS}
    for subs, subc, subt in (((1, 2, 1), (5, 4, 3), (1, 1, 2)), ((6, 2, 1), (5, 4, 3), (1, 1, 2)),
                             ((11, 6, 1), (2, 4, 3), (1, 1, 2))):
        for n in kr.KERNEL_VARS:
            d = rng.standard_normal(int(np.prod(subc))).astype(np.float32)
            fio.grid3d_put(p, n, d, subs, subc, subt)
            sl = tuple(slice(s - 1, s + (c - 1) * t, t) for s, c, t in zip(subs[::-1], subc[::-1], subt[::-1]))
            full[n][sl] = d.reshape(subc[::-1])
    with netcdf_file(p, "r", mmap=False) as nc:
        assert nc.title == b"Finite frequency kernel"
        assert dict(nc.dimensions) == {"I": 12, "J": 10, "K": 6}
        for n in kr.KERNEL_VARS:
            assert nc.variables[n].dimensions == ("K", "J", "I")
            np.testing.assert_array_equal(nc.variables[n][:], full[n])
    with pytest.raises(ValueError, match="outside"):
        fio.grid3d_put(p, "phase_Vp", np.zeros(8, np.float32), (12, 1, 1), (2, 2, 2), (1, 1, 1))


def test_runner_flow_with_oracle_device_layer(tmp_path, monkeypatch):
    """The whole runner (files, blocks, events, rotation, filters, windows,
    block writes, appending) on random SGT records, with the oracle's FP32
    routines standing in for the device kernels: the kernel files equal the
    oracle pipeline on the whole sub-grid."""
    import ctypes as C
    import types

    import kernel_case as kc
    from fwant_amd import kernel
    from oracle.pyoracle import Lib, build_metric
    lib = Lib(32)
    D = C.POINTER(C.c_double)
    P = lambda a: a.ctypes.data_as(C.POINTER(lib.R))  # noqa: E731

    class HostArray:
        def __init__(self, a):
            self.a = np.ascontiguousarray(a, np.float32)

        @classmethod
        def from_host(cls, a):
            return cls(np.array(a, np.float32))

        def to_host(self):
            return self.a.copy()

    def stress2strain(E, lam, mu):
        _, knt, npt = E.a.shape
        lib.stress2strain(npt, knt, *[P(E.a[c]) for c in range(6)], P(lam.a), P(mu.a))
        return E

    def sgt2scatter(R, S, lam, mu, dt):
        _, knt, npt = R.a.shape
        A, B = np.zeros((knt, npt), np.float32), np.zeros((knt, npt), np.float32)
        lib.scatter1(npt, knt, dt, (C.POINTER(lib.R) * 6)(*[P(R.a[c]) for c in range(6)]),
                     (C.POINTER(lib.R) * 6)(*[P(S.a[c]) for c in range(6)]), P(lam.a), P(mu.a), P(A), P(B))
        return HostArray(A), HostArray(B)

    def scatter_rotate(X, Y, Z, rot):
        knt, npt = X.a.shape
        r = np.ascontiguousarray(np.asarray(rot, np.float32).ravel(order="F"))
        lib.scatter_rotate(npt, knt, P(X.a), P(Y.a), P(Z.a), P(r))

    def _filter(mode, w0, b, a):
        knt, npt = w0.a.shape
        b, a = np.ascontiguousarray(b, np.float64), np.ascontiguousarray(a, np.float64)
        y = np.zeros_like(w0.a)
        lib.series_filter(mode, npt, knt, len(b), b.ctypes.data_as(D), a.ctypes.data_as(D), P(w0.a), P(y))
        return HostArray(y)

    def scatter2kernel(V, U, Sa, Sb, nt1, nt2, dt, M0):
        knt, npt = Sa.a.shape
        out = [np.zeros(npt, np.float32) for _ in range(4)]
        lib.scatter2kernel(npt, knt, nt1, nt2, dt, M0, P(np.ascontiguousarray(V, np.float32)),
                           P(np.ascontiguousarray(U, np.float32)), P(Sa.a), P(Sb.a), *[P(o) for o in out])
        return [HostArray(o) for o in out]

    fake = types.SimpleNamespace(stress2strain=stress2strain, sgt2scatter=sgt2scatter,
                                 scatter_rotate=scatter_rotate, _filter=_filter, scatter2kernel=scatter2kernel,
                                 filter_series_host=kernel.filter_series_host)
    monkeypatch.setattr(kr, "kernel", fake)
    monkeypatch.setattr(kr, "DeviceArray", HostArray)
    root = str(tmp_path)
    m = kc.model()
    rec = kc.random_records(5)
    metric = {k: v.astype(np.float32) for k, v in build_metric(m).items()}
    syn, T = kc.write_case(root, m, rec, metric=metric)
    st = kr.run(os.path.join(root, "TomoKernel.conf"), "sta.list", log=lambda *a: None)
    want = kc.oracle_kernels(m, rec, syn, T)
    saved = kc.check_run(root, st, want)
    kc.check_append_run(root, saved)
