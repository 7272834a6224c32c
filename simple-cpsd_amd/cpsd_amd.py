"""ctypes front end of the MI355X SD step (libcpsd_hip.so) and of its host
input builder (libcpsd_host.so).

The compute path is the HIP library; this module only moves numpy arrays
across the C-ABI declared in include/cpsd_abi.h and include/cpsd_host.h.
Loading fails loudly when a library is missing: there is no CPU fallback.

Inputs mirror the reference driver (Cpsd-mpi-bamboo/ccpsdv3/main.c:116-320):
the half-complex grid layouts, masker/tg/vc/vg tables, per-kind vl/vnl/vnlnrm,
0-based katm and the 27 Ewald image cells.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_SO = os.path.join(HERE, "libcpsd_hip.so")
HOST_SO = os.path.join(HERE, "libcpsd_host.so")
DATA = os.path.join(HERE, "data")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_DPTR = C.POINTER(C.c_double)

_hip = None
_host = None

SUPPORTED_GRIDS = (8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128)

# every symbol include/cpsd_abi.h declares
ABI_SYMBOLS = ("cpsd", "cr_fft3b_", "rc_fft3f_", "cpsd_last_error", "cpsd_set_abort_on_error",
               "cpsd_dev_create", "cpsd_dev_set_orbitals", "cpsd_dev_set_ions", "cpsd_dev_get_ions",
               "cpsd_dev_run", "cpsd_dev_get",
               "cpsd_dev_lagrange_info", "cpsd_dev_sizes", "cpsd_dev_profile",
               "cpsd_dev_profile_read", "cpsd_dev_destroy", "cpsd_dev_debug_fetch", "cpsd_fft3d_batch",
               "cpsd_lda_device", "cpsd_erfc_device", "cpsd_rotate", "cpsd_dev_epilogue",
               "cpsd_syev_sorted", "cpsd_kbpp_device", "cpsd_dev_set_precision", "cpsd_dev_set_graphs",
               "cpsd_dev_run_start", "cpsd_dev_run_finish", "cpsd_fft3d_batch_noalloc",
               "cpsd_fft3d_batch_noalloc_max", "cpsd_dev_set_preconditioner", "cpsd_mlp_nparam",
               "cpsd_dev_set_learned_preconditioner", "cpsd_dev_precond_features", "cpsd_dev_stats_begin",
               "cpsd_dev_stats_end", "cpsd_dev_stats_read", "cpsd_mlp_loss_grad", "cpsd_mlp_train",
               "cpsd_mlp_predict", "cpsd_dev_set_refresh", "cpsd_dev_set_fallback", "cpsd_dev_fallback_info",
               "cpsd_batch_create", "cpsd_batch_run", "cpsd_batch_info", "cpsd_batch_destroy",
               "cpsd_dev_autotune")
HOST_SYMBOLS = ("cpsdh_setup", "cpsdh_read_psp", "cpsdh_kbpp", "cpsdh_makepsi",
                "cpsdh_random_orbitals", "cpsdh_gcdotc", "cpsdh_write_elc", "cpsdh_read_elc",
                "cpsdh_write_ionout", "cpsdh_write_vpp", "cpsdh_read_vpp")


class CpsdError(RuntimeError):
    pass


def build(quiet=True):
    """Compile both libraries in place (hipcc for gfx950, gcc for the host part)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def hip_lib():
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_SO):
            raise CpsdError("libcpsd_hip.so is not built (%s); run __graft_entry__.build()" % HIP_SO)
        L = C.CDLL(HIP_SO)
        L.cpsd_last_error.restype = C.c_char_p
        L.cpsd_set_abort_on_error.argtypes = [C.c_int]
        L.cpsd_set_abort_on_error(0)
        L.cpsd_dev_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, _ip, _dp, _ip,
                                      _dp, _dp, _dp, _ip, _dp, _dp, _dp, _dp, C.c_double, C.c_int,
                                      _dp, C.c_double, C.c_double, _dp]
        L.cpsd_dev_set_orbitals.argtypes = [C.c_void_p, _dp]
        L.cpsd_dev_set_ions.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.cpsd_dev_get_ions.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cpsd_dev_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp, _DPTR, _DPTR, _DPTR]
        L.cpsd_dev_get.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.cpsd_dev_lagrange_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), _DPTR]
        L.cpsd_dev_set_precision.argtypes = [C.c_void_p, C.c_int]
        L.cpsd_dev_set_graphs.argtypes = [C.c_void_p, C.c_int]
        L.cpsd_dev_set_preconditioner.argtypes = [C.c_void_p, C.c_void_p]
        L.cpsd_dev_set_refresh.argtypes = [C.c_void_p, C.c_int]
        L.cpsd_dev_autotune.argtypes = [C.c_void_p, C.c_int, _ip]
        L.cpsd_batch_create.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int]
        L.cpsd_batch_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        L.cpsd_batch_info.argtypes = [C.c_void_p, C.POINTER(C.c_long), C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.cpsd_batch_destroy.argtypes = [C.c_void_p]
        L.cpsd_batch_destroy.restype = None
        L.cpsd_dev_set_fallback.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.cpsd_dev_fallback_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]
        _fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.cpsd_dev_set_learned_preconditioner.argtypes = [C.c_void_p, _fp, C.c_double, C.c_double, C.c_double]
        L.cpsd_dev_precond_features.argtypes = [C.c_void_p, _fp]
        L.cpsd_dev_stats_begin.argtypes = [C.c_void_p, _dp]
        L.cpsd_dev_stats_end.argtypes = [C.c_void_p]
        L.cpsd_dev_stats_read.argtypes = [C.c_void_p, _dp, _dp]
        L.cpsd_mlp_loss_grad.argtypes = [_fp, _dp, _dp, C.c_long, _fp, _fp, _DPTR]
        L.cpsd_mlp_train.argtypes = [_fp, _dp, _dp, C.c_long, _fp, C.c_int, C.c_float, _dp]
        L.cpsd_mlp_predict.argtypes = [_fp, C.c_long, _fp, _fp]
        L.cpsd_dev_run_start.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]
        L.cpsd_dev_run_finish.argtypes = [C.c_void_p, _dp, _DPTR, _DPTR, _DPTR]
        L.cpsd_dev_sizes.argtypes = [C.c_void_p, _ip]
        L.cpsd_dev_profile.argtypes = [C.c_void_p, C.c_int]
        L.cpsd_dev_profile_read.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _dp, _ip, C.c_int]
        L.cpsd_dev_destroy.argtypes = [C.c_void_p]
        L.cpsd_dev_debug_fetch.argtypes = [C.c_void_p, C.c_int, _dp, C.c_long]
        L.cpsd_dev_debug_fetch.restype = C.c_long
        L.cpsd_fft3d_batch.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.cpsd_fft3d_batch_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.cpsd_fft3d_batch_noalloc_max.argtypes = [C.c_int]
        L.cpsd_fft3d_batch_noalloc_max.restype = C.c_long
        L.cpsd_lda_device.argtypes = [C.c_int, _dp, _dp, _dp]
        L.cpsd_erfc_device.argtypes = [C.c_int, _dp, _dp]
        L.cpsd_dev_epilogue.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.cpsd_syev_sorted.argtypes = [C.c_int, _dp, _dp]
        L.cpsd_kbpp_device.argtypes = [C.c_double, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_int,
                                       _dp, _dp, _dp, _dp]
        L.cpsd_rotate.restype = None
        L.cpsd_rotate.argtypes = [C.c_int, _ip, C.c_int, _dp, _dp, _dp, _dp]
        L.cpsd.restype = None
        L.cpsd.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.c_int, C.c_int, _dp,
                           C.c_double, C.c_double, C.c_double, _ip, _dp, _dp, C.c_int, _dp,
                           C.c_double, C.c_double, _ip, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp,
                           _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _DPTR, _DPTR, _DPTR]
        for name in ("cr_fft3b_", "rc_fft3f_"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_dp] + [C.POINTER(C.c_int)] * 7 + [_dp, C.POINTER(C.c_int)]
        _hip = L
    return _hip


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_SO):
            raise CpsdError("libcpsd_host.so is not built (%s)" % HOST_SO)
        L = C.CDLL(HOST_SO)
        L.cpsdh_setup.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, _ip, _dp, _DPTR, _dp, _dp,
                                  _dp, _ip, _dp, _dp, _dp, _dp, _DPTR, _DPTR]
        L.cpsdh_read_psp.argtypes = [C.c_char_p, C.c_char_p, _DPTR, _DPTR, C.POINTER(C.c_int), _dp,
                                     C.POINTER(C.c_int), _DPTR, _dp, _dp, _dp, C.c_int]
        L.cpsdh_kbpp.argtypes = [C.c_double, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, C.c_int, _dp,
                                 _dp, _dp, _dp]
        L.cpsdh_makepsi.argtypes = [C.c_int, C.c_int, _ip, _dp]
        L.cpsdh_random_orbitals.argtypes = [C.c_int, C.c_int, _ip, _dp, C.c_ulonglong, C.c_double, _dp]
        L.cpsdh_gcdotc.argtypes = [C.c_int, _dp, _dp]
        L.cpsdh_gcdotc.restype = C.c_double
        L.cpsdh_write_elc.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _dp]
        L.cpsdh_read_elc.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), _DPTR,
                                     C.POINTER(C.c_int), _ip, C.c_void_p, C.c_long]
        L.cpsdh_write_ionout.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), _ip, C.c_int, _ip,
                                         _dp]
        _host = L
    return _host


def _err():
    return hip_lib().cpsd_last_error().decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise CpsdError("%s failed: %s" % (what, _err()))


class Batch:
    """independent cells advanced together, one graph launch per call
    (cpsd_batch_* of include/cpsd_abi.h)"""

    def __init__(self, devs):
        self.devs = list(devs)
        arr = (C.c_void_p * len(self.devs))(*[d.h.value for d in self.devs])
        h = C.c_void_p()
        _check(hip_lib().cpsd_batch_create(C.byref(h), arr, len(self.devs)), "cpsd_batch_create")
        self.h = h

    def run(self, inner, dt=5.8, fmass=400000.0):
        n = len(self.devs)
        E = np.zeros(10 * n)
        for i, d in enumerate(self.devs):
            E[10 * i] = d.E[0]
        de, dc, dr = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(hip_lib().cpsd_batch_run(self.h, inner, dt, fmass, E, de, dc, dr), "cpsd_batch_run")
        out = []
        for i, d in enumerate(self.devs):
            d.E[:] = E[10 * i:10 * i + 10]
            d.deltar = dr[i]
            out.append((d.E.copy(), float(de[i]), float(dc[i])))
        return out

    def info(self):
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        _check(hip_lib().cpsd_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)), "cpsd_batch_info")
        return dict(graph_launches=a.value, captures=b.value, eager=c.value)

    def close(self):
        if self.h:
            hip_lib().cpsd_batch_destroy(self.h)
            self.h = None


def run_batch(devs, inner, dt=5.8, fmass=400000.0):
    """one call of each of several independent cells: every call is started
    (each context on its own stream) before any is finished, so the cells
    advance concurrently; returns run()'s result per cell"""
    for d in devs:
        d.run_start(inner, dt, fmass)
    return [d.run_finish() for d in devs]


# ------------------------------------------------------------- restart files

def write_elc(path, icube, nfft, unit, c, ne):
    """ELCOUT/ELCIN file (ccpsdv3/main.c:539-547) for ispin = 1."""
    nea = np.array([ne, 0], dtype=np.int32)
    c = np.ascontiguousarray(c, dtype=np.float64)
    if c.size != ne * (nfft + 2) * nfft * nfft:
        raise CpsdError("write_elc: %d doubles for %d orbitals at %d^3" % (c.size, ne, nfft))
    if host_lib().cpsdh_write_elc(os.fsencode(path), icube, nfft, unit, 1, nea, c) != 0:
        raise CpsdError("cannot write %s" % path)


def read_elc(path):
    """Header and orbitals of an ELCIN/ELCOUT file (ccpsdv3/main.c:115-140)."""
    L = host_lib()
    ic, nf, isp = C.c_int(), C.c_int(), C.c_int()
    un = C.c_double()
    ne = np.zeros(2, dtype=np.int32)
    if L.cpsdh_read_elc(os.fsencode(path), C.byref(ic), C.byref(nf), C.byref(un), C.byref(isp), ne,
                        None, 0) != 0:
        raise CpsdError("cannot read the header of %s" % path)
    n = int(ne[0] + ne[1]) * (nf.value + 2) * nf.value * nf.value
    c = np.zeros(n)
    if L.cpsdh_read_elc(os.fsencode(path), C.byref(ic), C.byref(nf), C.byref(un), C.byref(isp), ne,
                        c.ctypes.data_as(C.c_void_p), n) != 0:
        raise CpsdError("cannot read the orbitals of %s" % path)
    return dict(icube=ic.value, nfft=nf.value, unit=un.value, ispin=isp.value, ne=ne.tolist(), c=c)


def write_ionout(path, fnames, lmax, katm, r2):
    """IONOUT (ccpsdv3/main.c:551-570)."""
    names = (C.c_char_p * len(fnames))(*[n.encode() for n in fnames])
    lm = np.ascontiguousarray(lmax, dtype=np.int32)
    ka = np.ascontiguousarray(katm, dtype=np.int32)
    r = np.ascontiguousarray(r2, dtype=np.float64)
    if host_lib().cpsdh_write_ionout(os.fsencode(path), len(fnames), names, lm, ka.size, ka, r) != 0:
        raise CpsdError("cannot write %s" % path)


# ------------------------------------------------------------------ inputs

def read_psp(path):
    L = host_lib()
    cap = 4096
    atom = C.create_string_buffer(8)
    zv, amass, drho = C.c_double(), C.c_double(), C.c_double()
    lmax, nrho = C.c_int(), C.c_int()
    rc = np.zeros(8)
    rho = np.zeros(cap)
    vp = np.zeros(4 * cap)
    wp = np.zeros(4 * cap)
    rc_ = L.cpsdh_read_psp(path.encode(), atom, C.byref(zv), C.byref(amass), C.byref(lmax), rc,
                           C.byref(nrho), C.byref(drho), rho, vp, wp, cap)
    if rc_ != 0:
        raise CpsdError("cannot parse %s (code %d)" % (path, rc_))
    n, lm = nrho.value, lmax.value
    return dict(atom=atom.value.decode(), zv=zv.value, amass=amass.value, lmax=lm,
                rc=rc[: lm + 1].copy(), nrho=n, drho=drho.value, rho=rho[:n].copy(),
                vp=vp[: (lm + 1) * n].copy(), wp=wp[: (lm + 1) * n].copy())


class Cell:
    """One periodic cell: setup tables (ccpsdv3/main.c:145-320) and the KB
    pseudopotential of every kind on its grid (ckbppv3/dkbppv3.c)."""

    def __init__(self, icube, unit, nfft, kinds, katm, positions, kb_device=False):
        """kinds: list of (psp path, lmax); katm: 0-based kind per ion;
        positions: (nion, 3) bohr."""
        L = host_lib()
        self.icube, self.unit, self.nfft = int(icube), float(unit), int(nfft)
        n = self.nfft
        self.nfft3d = (n // 2 + 1) * n * n
        self.n2ft3d = (n + 2) * n * n
        self.katm = np.ascontiguousarray(katm, dtype=np.int32)
        self.nion = len(self.katm)
        self.r = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1))
        self.psp = [read_psp(p) for p, _ in kinds]
        self.lmax = [min(l, p["lmax"]) for (_, l), p in zip(kinds, self.psp)]
        self.nkind = len(kinds)
        self.zv = np.array([p["zv"] for p in self.psp])
        self.amass = np.array([p["amass"] * 1822.89 for p in self.psp])
        self.lmmax = np.array([l * l for l in self.lmax], dtype=np.int32)
        om, cew, rcut = C.c_double(), C.c_double(), C.c_double()
        self.unita = np.zeros(9)
        self.unitg = np.zeros(9)
        self.G = np.zeros(3 * self.nfft3d)
        self.masker = np.zeros(self.n2ft3d, dtype=np.int32)
        self.tg = np.zeros(self.nfft3d)
        self.vc = np.zeros(self.nfft3d)
        self.vg = np.zeros(self.nfft3d)
        self.rcell = np.zeros(81)
        rc = L.cpsdh_setup(self.icube, self.unit, n, self.nion, self.katm, self.zv, C.byref(om),
                           self.unita, self.unitg, self.G, self.masker, self.tg, self.vc, self.vg,
                           self.rcell, C.byref(cew), C.byref(rcut))
        if rc != 0:
            raise CpsdError("cpsdh_setup failed (%d)" % rc)
        self.omega, self.cewald, self.rcut = om.value, cew.value, rcut.value
        self.vl = np.zeros(self.nkind * self.nfft3d)
        self.vnl = np.zeros(self.nkind * 9 * self.nfft3d)
        self.vnlnrm = np.zeros(self.nkind * 9)
        for k, (p, lm) in enumerate(zip(self.psp, self.lmax)):
            vl = np.zeros(self.nfft3d)
            vnl = np.zeros(9 * self.nfft3d)
            nrm = np.zeros(9)
            kb = hip_lib().cpsd_kbpp_device if kb_device else L.cpsdh_kbpp
            rc = kb(p["zv"], lm, p["nrho"], p["drho"], p["rho"], p["vp"], p["wp"], self.nfft3d, self.G,
                    vl, vnl, nrm)
            if rc != 0:
                raise CpsdError("KB transform failed" + (": " + _err() if kb_device else ""))
            self.vl[k * self.nfft3d:(k + 1) * self.nfft3d] = vl
            self.vnl[k * 9 * self.nfft3d:(k + 1) * 9 * self.nfft3d] = vnl
            self.vnlnrm[9 * k:9 * k + 9] = nrm
        self.ng = int(self.masker[0::2].sum())
        self.nproj = int(sum(self.lmmax[k] for k in self.katm))

    def plane_wave_orbitals(self, ne, psi0=None):
        spec = read_psi0(psi0 or os.path.join(DATA, "PSI0.DAT"), ne)
        psi = np.zeros(ne * self.n2ft3d)
        host_lib().cpsdh_makepsi(self.nfft, ne, spec, psi)
        return psi

    def random_orbitals(self, ne, seed=1, decay=0.0):
        psi = np.zeros(ne * self.n2ft3d)
        rc = host_lib().cpsdh_random_orbitals(self.nfft, ne, self.masker, self.G, seed, decay, psi)
        if rc != 0:
            raise CpsdError("cpsdh_random_orbitals failed")
        return psi


def read_psi0(path, ne):
    rows = []
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) >= 4:
                rows.append([int(x) for x in t[:4]])
            if len(rows) == ne:
                break
    if len(rows) < ne:
        raise CpsdError("%s holds only %d orbitals" % (path, len(rows)))
    return np.ascontiguousarray(np.array(rows, dtype=np.int32).reshape(-1))


# water geometry used by the workloads: O-H 1.81 bohr, H-O-H 104.5 degrees
OH_BOHR = 1.81
HOH_DEG = 104.5


def water_molecule(center=(0.0, 0.0, 0.0), rot=None):
    half = np.deg2rad(HOH_DEG) / 2
    h1 = np.array([OH_BOHR * np.sin(half), OH_BOHR * np.cos(half), 0.0])
    h2 = np.array([-OH_BOHR * np.sin(half), OH_BOHR * np.cos(half), 0.0])
    pts = np.array([np.zeros(3), h1, h2])
    if rot is not None:
        pts = pts @ rot.T
    return pts + np.asarray(center)


def water_box(nmol, L, seed=0, jitter=0.3):
    """nmol waters on a near-cubic lattice filling an L bohr cube, randomly
    oriented, with +-jitter bohr uniform displacement of each molecule
    (seeded numpy Generator).  Returns katm (O=0, H=1) and positions."""
    rng = np.random.default_rng(seed)
    nx = int(np.ceil(nmol ** (1.0 / 3.0)))
    dims = [nx, nx, nx]
    while (dims[0] * dims[1] * (dims[2] - 1)) >= nmol:
        dims[2] -= 1
    sites = [(i, j, k) for k in range(dims[2]) for j in range(dims[1]) for i in range(dims[0])][:nmol]
    katm, pos = [], []
    for (i, j, k) in sites:
        c = (np.array([i, j, k]) + 0.5) * (L / np.array(dims)) - L / 2
        c = c + rng.uniform(-jitter, jitter, 3)
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        m = water_molecule(c, q)
        katm += [0, 1, 1]
        pos.append(m)
    return np.array(katm, dtype=np.int32), np.concatenate(pos)


def water_batch_specs(seed, count=64):
    """(nmol, L, nfft) of `count` seeded water cells of 1-8 molecules: cell
    edge 12-22 bohr growing with the molecule count, grid 24^3-64^3 by edge;
    the cells of the batched workload carry 4 orbitals per molecule"""
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(count):
        nmol = int(rng.integers(1, 9))
        L = 12.0 + 10.0 * (nmol - 1) / 7.0 + float(rng.uniform(-0.4, 0.4))
        n = 24 if L < 14 else 32 if L < 17 else 48 if L < 20 else 64
        cells.append((nmol, L, n))
    return cells


def water_cell(nmol, L, nfft, seed=0, jitter=0.3, psp_dir=None, kb_device=False):
    """nmol waters (O s+p, H s projectors) in an L-bohr cubic cell; kb_device
    runs the KB transform on the GPU (cpsd_kbpp_device) instead of the host."""
    d = psp_dir or os.path.join(DATA, "psp")
    katm, pos = water_box(nmol, L, seed, jitter)
    return Cell(1, L, nfft, [(os.path.join(d, "O00.dat"), 2), (os.path.join(d, "H00.dat"), 1)],
                katm, pos, kb_device=kb_device)


# --------------------------------------------------------- device SD step

class DeviceSD:
    """Orbitals resident in HBM across calls (cpsd_dev_* of include/cpsd_abi.h)."""

    def __init__(self, cell, ne):
        self.cell, self.ne = cell, int(ne)
        h = C.c_void_p()
        c = cell
        _check(hip_lib().cpsd_dev_create(C.byref(h), c.nfft, self.ne, c.nion, c.katm, c.zv, c.lmmax,
                                         c.vnlnrm, c.vl, c.vnl, c.masker, c.tg, c.vc, c.vg, c.unitg,
                                         c.omega, 27, c.rcell, c.cewald, c.rcut, c.r),
               "cpsd_dev_create")
        self.h = h
        self.E = np.zeros(10)

    def set_orbitals(self, psi):
        _check(hip_lib().cpsd_dev_set_orbitals(self.h, np.ascontiguousarray(psi, dtype=np.float64)),
               "cpsd_dev_set_orbitals")

    def run_start(self, inner, dt=5.8, fmass=400000.0):
        """enqueue a call and return at once (see run_batch)"""
        _check(hip_lib().cpsd_dev_run_start(self.h, inner, dt, fmass, self.E), "cpsd_dev_run_start")

    def run_finish(self):
        de, dc, dr = C.c_double(), C.c_double(), C.c_double()
        _check(hip_lib().cpsd_dev_run_finish(self.h, self.E, C.byref(de), C.byref(dc), C.byref(dr)),
               "cpsd_dev_run_finish")
        self.deltar = dr.value
        return self.E.copy(), de.value, dc.value

    def set_ions(self, move, r2=None, amass=None):
        """Positions (default: the cell's) and, with move=1, ion motion with
        the cell's masses."""
        c = self.cell
        r2 = np.ascontiguousarray(c.r if r2 is None else r2, dtype=np.float64)
        am = np.ascontiguousarray(c.amass if amass is None else amass, dtype=np.float64)
        _check(hip_lib().cpsd_dev_set_ions(self.h, int(move), am, r2), "cpsd_dev_set_ions")

    def get_ions(self):
        n = 3 * self.cell.nion
        r1, r2, f = np.zeros(n), np.zeros(n), np.zeros(n)
        _check(hip_lib().cpsd_dev_get_ions(self.h, r1.ctypes.data_as(C.c_void_p),
                                           r2.ctypes.data_as(C.c_void_p),
                                           f.ctypes.data_as(C.c_void_p)), "cpsd_dev_get_ions")
        return dict(r1=r1, r2=r2, fion=f)

    def run(self, inner, dt=5.8, fmass=400000.0):
        de, dc, dr = C.c_double(), C.c_double(), C.c_double()
        _check(hip_lib().cpsd_dev_run(self.h, inner, dt, fmass, self.E, C.byref(de), C.byref(dc),
                                      C.byref(dr)), "cpsd_dev_run")
        self.deltar = dr.value
        return self.E.copy(), de.value, dc.value

    def get(self, c1=False, c2=False, cpsi=False, dn=False, hml=False, lmd=False):
        c = self.cell
        out = {}
        arrs = []
        for name, want, size in (("c1", c1, self.ne * c.n2ft3d), ("c2", c2, self.ne * c.n2ft3d),
                                 ("cpsi", cpsi, self.ne * c.n2ft3d), ("dn", dn, c.n2ft3d),
                                 ("hml", hml, self.ne * self.ne), ("lmd", lmd, self.ne * self.ne)):
            if want:
                out[name] = np.zeros(size)
                arrs.append(out[name].ctypes.data_as(C.c_void_p))
            else:
                arrs.append(None)
        _check(hip_lib().cpsd_dev_get(self.h, *arrs), "cpsd_dev_get")
        return out

    def epilogue(self):
        """The run epilogue (ccpsdv3/main.c:451-484) on the resident state:
        orthonormality matrix of c1, eigenpairs of hml in eigsrt order, and
        the device's c1 replaced by c2 V."""
        ne = self.ne
        ortho, eig, V = np.zeros(ne * ne), np.zeros(ne), np.zeros(ne * ne)
        _check(hip_lib().cpsd_dev_epilogue(self.h, ortho.ctypes.data_as(C.c_void_p),
                                           eig.ctypes.data_as(C.c_void_p),
                                           V.ctypes.data_as(C.c_void_p)), "cpsd_dev_epilogue")
        return dict(ortho=ortho.reshape(ne, ne), eig=eig, V=V)

    def set_precision(self, mode):
        """0: FP64 throughout; 1: FP32 orbital FFTs and KB products in every
        step of a run() but its last (energies stay FP64 values of the current
        orbitals); 2: as 1 with a BF16-input KB projection; 3: FP32 FFTs and
        every product of those steps (KB projection and apply, preconditioned
        update, overlaps of the step c2 - c1, constraint) in split BF16
        (ne 128 or 256)"""
        _check(hip_lib().cpsd_dev_set_precision(self.h, int(mode)), "cpsd_dev_set_precision")

    def set_preconditioner(self, K):
        """K: one factor per half-complex G ([nfft3d], the reference's layout)
        for the preconditioned step c2 = c1 + K o dte (cpsi + c1 hml), or None
        for the reference's step"""
        if K is None:
            _check(hip_lib().cpsd_dev_set_preconditioner(self.h, None), "cpsd_dev_set_preconditioner")
            return
        K = np.ascontiguousarray(K, dtype=np.float64)
        if K.size != self.cell.nfft3d:
            raise CpsdError("preconditioner needs %d factors, got %d" % (self.cell.nfft3d, K.size))
        _check(hip_lib().cpsd_dev_set_preconditioner(self.h, K.ctypes.data), "cpsd_dev_set_preconditioner")

    def autotune(self, steps=3):
        """measure and set the fastest launch plan of this size class (the
        orbitals are kept, the run state restarts)"""
        p = np.zeros(5, dtype=np.int32)
        _check(hip_lib().cpsd_dev_autotune(self.h, int(steps), p), "cpsd_dev_autotune")
        return dict(fft_batch=int(p[0]), density_plane_orbitals=(int(p[1]), int(p[2])), splitk_target=int(p[3]),
                    from_table=bool(p[4]))

    def set_refresh(self, period):
        """FP64 refresh every `period` steps of the mixed modes (0: last step of each call)"""
        _check(hip_lib().cpsd_dev_set_refresh(self.h, int(period)), "cpsd_dev_set_refresh")

    def set_fallback(self, enable=True, rise=1e-8):
        """fall back to the FP64 step when a call's energy exceeds the previous
        call's by more than rise * max(1, |E|)"""
        _check(hip_lib().cpsd_dev_set_fallback(self.h, 1 if enable else 0, float(rise)), "cpsd_dev_set_fallback")

    def fallback_info(self):
        f, m, p = C.c_int(), C.c_int(), C.c_int()
        _check(hip_lib().cpsd_dev_fallback_info(self.h, C.byref(f), C.byref(m), C.byref(p)),
               "cpsd_dev_fallback_info")
        return dict(fallbacks=f.value, mode=m.value, precond=p.value)

    def set_learned_preconditioner(self, params, alpha, kmin=0.0, kmax=1e3):
        """K = clamp(alpha exp(MLP(features)), kmin, kmax) on this cell's G"""
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != MLP_NPARAM:
            raise CpsdError("the network has %d parameters, got %d" % (MLP_NPARAM, p.size))
        _check(hip_lib().cpsd_dev_set_learned_preconditioner(self.h, p, alpha, kmin, kmax),
               "cpsd_dev_set_learned_preconditioner")

    def precond_features(self):
        X = np.zeros((self.sizes()["ng"], 4), dtype=np.float32)
        _check(hip_lib().cpsd_dev_precond_features(self.h, X.reshape(-1)), "cpsd_dev_precond_features")
        return X

    def stats_begin(self, cinf):
        _check(hip_lib().cpsd_dev_stats_begin(self.h, np.ascontiguousarray(cinf, dtype=np.float64)),
               "cpsd_dev_stats_begin")

    def stats_end(self):
        _check(hip_lib().cpsd_dev_stats_end(self.h), "cpsd_dev_stats_end")

    def stats_read(self):
        ng = self.sizes()["ng"]
        A, B = np.zeros(ng), np.zeros(ng)
        _check(hip_lib().cpsd_dev_stats_read(self.h, A, B), "cpsd_dev_stats_read")
        return A, B

    def set_graphs(self, enable=True):
        """capture each run() into a HIP graph and replay it on later calls
        with the same arguments and state parity"""
        _check(hip_lib().cpsd_dev_set_graphs(self.h, 1 if enable else 0), "cpsd_dev_set_graphs")

    def lagrange_info(self):
        it, ad = C.c_int(), C.c_double()
        hip_lib().cpsd_dev_lagrange_info(self.h, C.byref(it), C.byref(ad))
        return it.value, ad.value

    def sizes(self):
        s = np.zeros(3, dtype=np.int32)
        hip_lib().cpsd_dev_sizes(self.h, s)
        return dict(ng=int(s[0]), ngp=int(s[1]), nproj=int(s[2]))

    def profile(self, enable=True):
        hip_lib().cpsd_dev_profile(self.h, 1 if enable else 0)

    def profile_read(self):
        names = C.create_string_buffer(1024)
        ms = np.zeros(32)
        calls = np.zeros(32, dtype=np.int32)
        k = hip_lib().cpsd_dev_profile_read(self.h, names, 1024, ms, calls, 32)
        nm = names.value.decode().split(";")
        return {nm[i]: (float(ms[i]), int(calls[i])) for i in range(k)}

    def fetch(self, what):
        cap = 1 << 28
        buf = np.zeros(min(cap, 4 * self.ne * self.cell.n2ft3d + 64))
        n = hip_lib().cpsd_dev_debug_fetch(self.h, what, buf, buf.size)
        if n < 0:
            raise CpsdError(_err())
        return buf[:n].copy()

    def close(self):
        if self.h:
            hip_lib().cpsd_dev_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cpsd_dropin(cell, ne, inner, c2, E, dt=5.8, fmass=400000.0, move=0, r2=None):
    """Call the drop-in cpsd() symbol exactly as ccpsdv3/main.c:422-426 does.
    Returns dict of the arrays cpsd() writes."""
    c = cell
    L = hip_lib()
    n2 = c.n2ft3d
    nea = np.array([ne, 0], dtype=np.int32)
    r1 = np.zeros(3 * c.nion)
    r2 = (c.r if r2 is None else np.asarray(r2, dtype=np.float64)).copy()
    fion = np.zeros(3 * c.nion)
    c1 = np.zeros(ne * n2)
    c2 = np.ascontiguousarray(c2, dtype=np.float64).copy()
    cpsi = np.zeros(ne * n2)
    dn = np.zeros(n2)
    hml = np.zeros(ne * ne)
    lmd = np.zeros(ne * ne)
    work = np.zeros(7 * n2 + c.nion * (5 * c.nfft + 3) + 7 * ne * ne)
    E = np.ascontiguousarray(E, dtype=np.float64).copy()
    de, dc, dr = C.c_double(), C.c_double(), C.c_double()
    L.cpsd(inner, move, c.nion, 1, nea, c.nfft, c.nfft3d, n2, c.unitg, c.omega, dt, fmass, c.katm,
           c.amass, c.zv, 27, c.rcell, c.cewald, c.rcut, c.masker, c.G, c.tg, c.vc, c.vg,
           c.lmmax.copy(), c.vnlnrm.copy(), c.vl, c.vnl, r1, r2, fion, c1, c2, cpsi, dn, hml, lmd,
           work, E, C.byref(de), C.byref(dc), C.byref(dr))
    return dict(c1=c1, c2=c2, cpsi=cpsi, dn=dn, hml=hml, lmd=lmd, E=E, deltae=de.value,
                deltac=dc.value, deltar=dr.value, r1=r1, r2=r2, fion=fion)


def rotate_dropin(ne, n2ft3d, hml, c2):
    """The drop-in cpsd_rotate() (ccpsdv3/main.c:466-484) after a cpsd_dropin()
    call: returns the eigenvectors (hml layout), eigenvalues and c1 = c2 V."""
    h = np.ascontiguousarray(hml, dtype=np.float64).copy()
    eig = np.zeros(ne)
    c1 = np.zeros(ne * n2ft3d)
    nea = np.array([ne, 0], dtype=np.int32)
    hip_lib().cpsd_rotate(1, nea, n2ft3d, h, eig, np.ascontiguousarray(c2, dtype=np.float64), c1)
    return dict(V=h, eig=eig, c1=c1)


def syev_sorted(a):
    """Eigenpairs of a symmetric matrix (upper triangle of the column-major
    array read) in eigsrt order: (eigenvalues, eigenvectors as columns of a
    column-major n*n array)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = int(round(np.sqrt(a.size)))
    v = a.reshape(-1).copy()
    eig = np.zeros(n)
    _check(hip_lib().cpsd_syev_sorted(n, v, eig), "cpsd_syev_sorted")
    return eig, v


MLP_NPARAM = 1249


def mlp_loss_grad(X, A, B, params):
    """loss L = sum A K^2 - 2 B K (K = exp(MLP(X))) and dL/dparams, on the GPU"""
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1)
    g = np.zeros(MLP_NPARAM, dtype=np.float32)
    loss = C.c_double()
    _check(hip_lib().cpsd_mlp_loss_grad(X, np.ascontiguousarray(A, dtype=np.float64),
                                        np.ascontiguousarray(B, dtype=np.float64), X.size // 4,
                                        np.ascontiguousarray(params, dtype=np.float32), g, C.byref(loss)),
           "cpsd_mlp_loss_grad")
    return loss.value, g


def mlp_train(X, A, B, params, iters, lr):
    """iters Adam steps on the GPU; returns (params, loss history)"""
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1)
    p = np.array(params, dtype=np.float32)
    hist = np.zeros(iters)
    _check(hip_lib().cpsd_mlp_train(X, np.ascontiguousarray(A, dtype=np.float64),
                                    np.ascontiguousarray(B, dtype=np.float64), X.size // 4, p, int(iters),
                                    float(lr), hist), "cpsd_mlp_train")
    return p, hist


def mlp_predict(X, params):
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1)
    y = np.zeros(X.size // 4, dtype=np.float32)
    _check(hip_lib().cpsd_mlp_predict(X, X.size // 4, np.ascontiguousarray(params, dtype=np.float32), y),
           "cpsd_mlp_predict")
    return y


def fft3d(direction, nfft, a):
    """Batched in-place FFT of reference-layout arrays: direction +1 is
    cr_fft3b, -1 is rc_fft3f.  a: (nbatch, (N+2) N N)."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    nb = a.size // ((nfft + 2) * nfft * nfft)
    _check(hip_lib().cpsd_fft3d_batch(direction, nfft, nb, a.reshape(-1)), "cpsd_fft3d_batch")
    return a


def fortran_fft(direction, nfft, a):
    """One transform through the Fortran-ABI symbols cr_fft3b_/rc_fft3f_."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    n = nfft
    ints = [C.c_int(v) for v in (n // 2 + 1, (n // 2 + 1) * n, n + 2, (n + 2) * n, n, n, n)]
    tmp = np.zeros(2 * (n * n + 10) + 40 * n)
    ntmp = C.c_int(tmp.size)
    f = hip_lib().cr_fft3b_ if direction > 0 else hip_lib().rc_fft3f_
    f(a, *[C.byref(v) for v in ints], tmp, C.byref(ntmp))
    return a


def lda_device(dn):
    dn = np.ascontiguousarray(dn, dtype=np.float64)
    p = np.zeros_like(dn)
    e = np.zeros_like(dn)
    _check(hip_lib().cpsd_lda_device(dn.size, dn, p, e), "cpsd_lda_device")
    return p, e


def erfc_device(w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros_like(w)
    _check(hip_lib().cpsd_erfc_device(w.size, w, out), "cpsd_erfc_device")
    return out
