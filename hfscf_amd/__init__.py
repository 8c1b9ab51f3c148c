"""hfscf_amd — MI355X-native RHF two-electron engine behind hfscf's seams.

Python bindings over two C ABIs:
  * ``include/hfgpu.h``  (lib/libhfgpu.so)  the GPU engine: Schwarz screening, shell-quartet
    ERIs, direct J/K / Fock builds and the adjoint two-electron gradient (gfx950 HIP kernels);
  * ``include/hfscf.h``  (lib/libhfscf.so)  the host RHF layer: basis and molecule handling,
    one-electron integrals, DIIS, the SCF loop and gradient assembly.

The names follow hfscf: ``Molecule`` (MOLEC::Molecule), ``rhf`` (Mol::hf_solver::scf_run),
``rhf_gradient`` (Mol::scf_gradient::calc_scf_gradient_rhf), ``GpuContext`` (the Erios /
update_fock_matrix replacement). There is no CPU fallback: without the built libraries or a
GPU every call raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
BASISDIR = os.path.join(HERE, "basis")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

FP64, MIXED = 0, 1


class HfgOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("integral_tol", C.c_double), ("schwarz", C.c_int),
                ("density_tol", C.c_double), ("graph", C.c_int), ("work_split", C.c_int),
                ("coop_min_l", C.c_int), ("side_streams", C.c_int), ("group_shells", C.c_int),
                ("reserved", C.c_int * 1)]


class HfgStats(C.Structure):
    _fields_ = [("quartets", C.c_longlong), ("model_flops", C.c_double), ("kernel_seconds", C.c_double),
                ("total_seconds", C.c_double), ("work_items", C.c_longlong), ("slab_bytes", C.c_longlong)]


class ScfOptions(C.Structure):
    _fields_ = [("backend", C.c_int), ("device", C.c_int), ("precision", C.c_int), ("incremental", C.c_int),
                ("diis_range", C.c_int), ("max_iter", C.c_int), ("schwarz", C.c_int), ("verbose", C.c_int),
                ("energy_tol", C.c_double), ("rms_tol", C.c_double), ("integral_tol", C.c_double),
                ("final_fp64", C.c_int), ("density_tol", C.c_double), ("graph", C.c_int), ("autotune", C.c_int),
                ("rebuild_every", C.c_int), ("device_step", C.c_int), ("polish_rms", C.c_double)]


class ScfResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("energy_elec", C.c_double), ("energy_nuc", C.c_double),
                ("iterations", C.c_int), ("converged", C.c_int), ("fock_seconds", C.c_double),
                ("last_fock_seconds", C.c_double), ("fock_builds", C.c_int), ("energy_final", C.c_double)]


class BatchStats(C.Structure):
    _fields_ = [("setup_seconds", C.c_double), ("scf_seconds", C.c_double), ("fock_seconds", C.c_double),
                ("gradient_seconds", C.c_double), ("autotune_seconds", C.c_double), ("iterations", C.c_int),
                ("work_items", C.c_int), ("quartets", C.c_longlong)]


JKFUNC = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _dp, _dp)
GRADFUNC = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, _dp)


class Callbacks(C.Structure):
    _fields_ = [("user", C.c_void_p), ("build_jk", JKFUNC), ("grad2e", GRADFUNC)]


_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(LIBDIR, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run __graft_entry__.build() (make -C hfscf_amd/csrc)")
        _libs[name] = C.CDLL(path)
    return _libs[name]


def hfgpu():
    """ctypes handle of libhfgpu.so with argument types declared."""
    lib = _load("libhfgpu.so")
    if not getattr(lib, "_typed", False):
        lib.hfg_default_options.argtypes = [C.POINTER(HfgOptions)]
        lib.hfg_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.POINTER(HfgOptions)]
        lib.hfg_set_basis.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp, C.c_int, _ip, _ip]
        lib.hfg_set_geometry.argtypes = [C.c_void_p, C.c_int, _dp, _ip]
        lib.hfg_build_jk.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int]
        lib.hfg_fock.argtypes = [C.c_void_p, _dp, _dp, C.c_int, _dp]
        lib.hfg_reset_incremental.argtypes = [C.c_void_p]
        lib.hfg_set_precision.argtypes = [C.c_void_p, C.c_int]
        lib.hfg_set_density_tol.argtypes = [C.c_void_p, C.c_double]
        lib.hfg_batch_add_molecule.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp, C.c_int, _ip, _ip,
                                               C.c_int, _dp]
        lib.hfg_batch_commit.argtypes = [C.c_void_p]
        lib.hfg_batch_layout.argtypes = [C.c_void_p, _ip, _ip, C.POINTER(C.c_longlong), _ip]
        lib.hfg_autotune.argtypes = [C.c_void_p, C.c_int]
        lib.hfg_class_stats.argtypes = [C.c_void_p, C.c_int, _ip, C.POINTER(C.c_longlong), _dp, _dp]
        lib.hfg_launch_shapes.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip, _ip]
        lib.hfg_energy.argtypes = [C.c_void_p, _dp, _dp]
        lib.hfg_energy_backward.argtypes = [C.c_void_p, _dp, _dp]
        lib.hfg_oneint.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.hfg_eri_vector.argtypes = [C.c_void_p, _dp, C.c_longlong]
        lib.hfg_hessian_2e.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.hfg_oneint_gradient.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.hfg_eri_shell_quartet.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        lib.hfg_schwarz.argtypes = [C.c_void_p, _dp]
        lib.hfg_copy_calibration.argtypes = [C.c_int, C.c_longlong, C.c_int, _dp]
        lib.hfg_boys_probe.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, C.c_int]
        lib.hfg_build_jk_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.hfg_last_stats.argtypes = [C.c_void_p, C.POINTER(HfgStats)]
        lib.hfg_nbf.argtypes = [C.c_void_p]
        lib.hfg_last_error.argtypes = [C.c_void_p]
        lib.hfg_last_error.restype = C.c_char_p
        lib.hfg_destroy.argtypes = [C.c_void_p]
        lib._typed = True
    return lib


def hfscf():
    """ctypes handle of libhfscf.so with argument types declared."""
    lib = _load("libhfscf.so")
    if not getattr(lib, "_typed", False):
        lib.hfscf_default_scf_options.argtypes = [C.POINTER(ScfOptions)]
        lib.hfscf_mol_create.argtypes = [C.c_int, _ip, _dp, C.c_char_p, C.c_int]
        lib.hfscf_mol_create.restype = C.c_void_p
        lib.hfscf_mol_from_xyz.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        lib.hfscf_mol_from_xyz.restype = C.c_void_p
        lib.hfscf_mol_destroy.argtypes = [C.c_void_p]
        lib.hfscf_mol_info.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _ip]
        lib.hfscf_mol_nprim_total.argtypes = [C.c_void_p]
        lib.hfscf_mol_shells.argtypes = [C.c_void_p, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp]
        lib.hfscf_mol_geometry.argtypes = [C.c_void_p, _ip, _dp]
        lib.hfscf_mol_set_geometry.argtypes = [C.c_void_p, _dp]
        lib.hfscf_mol_enuc.argtypes = [C.c_void_p]
        lib.hfscf_mol_enuc.restype = C.c_double
        lib.hfscf_oneint.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.hfscf_oneint_gradient.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.hfscf_rhf_run.argtypes = [C.c_void_p, C.POINTER(ScfOptions), C.POINTER(Callbacks),
                                      C.POINTER(ScfResult), _dp, _dp, _dp]
        lib.hfscf_rhf_gradient.argtypes = [C.c_void_p, C.POINTER(ScfOptions), C.POINTER(Callbacks), _dp, _dp, _dp]
        lib.hfscf_batch_rhf.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(ScfOptions), C.POINTER(ScfResult),
                                        _dp, _dp, _dp, C.POINTER(BatchStats)]
        lib.hfscf_last_error.restype = C.c_char_p
        lib.hfscf_input_read.argtypes = [C.c_char_p]
        lib.hfscf_input_read.restype = C.c_void_p
        lib.hfscf_input_destroy.argtypes = [C.c_void_p]
        lib.hfscf_input_info.argtypes = [C.c_void_p, _ip, _ip, _ip]
        lib.hfscf_input_atoms.argtypes = [C.c_void_p, _ip, _dp]
        lib.hfscf_input_setting.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
        lib.hfscf_input_basis_file.argtypes = [C.c_void_p]
        lib.hfscf_input_basis_file.restype = C.c_char_p
        lib.hfscf_input_comment.argtypes = [C.c_void_p]
        lib.hfscf_input_comment.restype = C.c_char_p
        lib._typed = True
    return lib


def read_input(path):
    """An hfscf .dat input file: comment, keywords (validated against hfscf's table), atoms (Z,
    bohr), charge, multiplicity and the basis file name (hfscf_input_read)."""
    lib = hfscf()
    h = lib.hfscf_input_read(str(path).encode())
    if not h:
        raise ValueError(lib.hfscf_last_error().decode())
    try:
        n, q, m = C.c_int(), C.c_int(), C.c_int()
        lib.hfscf_input_info(h, C.byref(n), C.byref(q), C.byref(m))
        Z = np.zeros(n.value, dtype=np.int32)
        X = np.zeros(3 * n.value)
        lib.hfscf_input_atoms(h, Z.ctypes.data_as(_ip), _ptr(X))
        settings = {}
        k, v = C.c_char_p(), C.c_char_p()
        for i in range(lib.hfscf_input_setting(h, 0, None, None)):
            lib.hfscf_input_setting(h, i, C.byref(k), C.byref(v))
            settings[k.value.decode()] = v.value.decode()
        return dict(comment=lib.hfscf_input_comment(h).decode(), settings=settings, Z=Z, xyz=X.reshape(-1, 3),
                    charge=q.value, multiplicity=m.value, basis_file=lib.hfscf_input_basis_file(h).decode())
    finally:
        lib.hfscf_input_destroy(h)


def _ptr(a, dtype=np.float64):
    return a.ctypes.data_as(_dp if dtype == np.float64 else _ip)


def basis_path(name):
    """hfscf's basis file naming: lowercase, '*' -> '_st_' (hfscf: src/molecule/hfscf_parser.hpp:134-151)."""
    fn = name.lower().replace("*", "_st_") + ".gbs"
    path = os.path.join(BASISDIR, fn)
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return path


class Molecule:
    """A molecule with its basis (MOLEC::Molecule)."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError(hfscf().hfscf_last_error().decode())
        self._h = handle
        lib = hfscf()
        vals = [C.c_int() for _ in range(6)]
        lib.hfscf_mol_info(self._h, *[C.byref(v) for v in vals])
        self.natom, self.nshell, self.nbf, self.ncart, self.nelec, pure = [v.value for v in vals]
        self.pure = bool(pure)

    @classmethod
    def from_xyz(cls, xyz_path, basis, pure=-1):
        return cls(hfscf().hfscf_mol_from_xyz(str(xyz_path).encode(), basis_path(basis).encode(), pure))

    @classmethod
    def from_arrays(cls, Z, xyz_bohr, basis, pure=-1):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        return cls(hfscf().hfscf_mol_create(len(Z), Z.ctypes.data_as(_ip), _ptr(x), basis_path(basis).encode(), pure))

    @classmethod
    def from_input(cls, path):
        """The molecule of an hfscf .dat input (its atoms, basis_set and use_puream)."""
        inp = read_input(path)
        if "basis_set" not in inp["settings"]:
            raise ValueError(f"{path}: no basis_set keyword")
        pure = {"true": 1, "false": 0}.get(inp["settings"].get("use_puream"), -1)
        return cls.from_arrays(inp["Z"], inp["xyz"], inp["settings"]["basis_set"], pure)

    def __del__(self):
        if getattr(self, "_h", None):
            hfscf().hfscf_mol_destroy(self._h)
            self._h = None

    def shells(self):
        """Shell table: dict of arrays (L, atom, nprim, cart_off, pure_off, alpha, coef, coef_raw)."""
        lib = hfscf()
        ns, npt = self.nshell, lib.hfscf_mol_nprim_total(self._h)
        ints = [np.zeros(ns, dtype=np.int32) for _ in range(5)]
        dbl = [np.zeros(npt) for _ in range(3)]
        lib.hfscf_mol_shells(self._h, *[a.ctypes.data_as(_ip) for a in ints], *[_ptr(a) for a in dbl])
        keys = ["L", "atom", "nprim", "cart_off", "pure_off"]
        out = dict(zip(keys, ints))
        out.update(alpha=dbl[0], coef=dbl[1], coef_raw=dbl[2])
        return out

    def geometry(self):
        Z = np.zeros(self.natom, dtype=np.int32)
        x = np.zeros(3 * self.natom)
        hfscf().hfscf_mol_geometry(self._h, Z.ctypes.data_as(_ip), _ptr(x))
        return Z, x.reshape(-1, 3)

    def set_geometry(self, xyz_bohr):
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        hfscf().hfscf_mol_set_geometry(self._h, _ptr(x))

    def enuc(self):
        return hfscf().hfscf_mol_enuc(self._h)

    def oneint(self):
        n = self.nbf
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        if hfscf().hfscf_oneint(self._h, _ptr(S), _ptr(T), _ptr(V)):
            raise RuntimeError(hfscf().hfscf_last_error().decode())
        return S, T, V

    def oneint_gradient(self, D, Q):
        g = np.zeros((self.natom, 3))
        D = np.ascontiguousarray(D)
        Q = np.ascontiguousarray(Q)
        hfscf().hfscf_oneint_gradient(self._h, _ptr(D), _ptr(Q), _ptr(g))
        return g


ANGSTROM_TO_BOHR = 1.8897261254535  # hfscf: src/molecule/hfscf_constants.hpp:5
_SYMBOLS = {"H": 1, "HE": 2, "LI": 3, "BE": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "NE": 10}


def copy_calibration(mib, reps=3, device=0):
    """Runs `reps` launches of a plain copy kernel of `mib` MiB (reads and writes that many bytes
    each) through libhfgpu; returns the milliseconds per copy. A kernel of known HBM traffic
    for scaling a profiler's byte counters."""
    ms = C.c_double(0.0)
    rc = hfgpu().hfg_copy_calibration(int(device), int(mib * 2 ** 20), int(reps), C.byref(ms))
    if rc:
        raise RuntimeError(f"hfg_copy_calibration failed ({rc})")
    return ms.value


def read_xyz_frames(path):
    """Frames of a multi-frame .xyz file (angstrom) as (Z, xyz_bohr) pairs."""
    frames = []
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines):
        if not lines[k].strip():
            k += 1
            continue
        nat = int(lines[k])
        Z, X = [], []
        for line in lines[k + 2:k + 2 + nat]:
            sym, x, y, z = line.split()[:4]
            Z.append(_SYMBOLS[sym.upper()])
            X.append([float(x), float(y), float(z)])
        frames.append((np.array(Z, dtype=np.int32), np.array(X) * ANGSTROM_TO_BOHR))
        k += 2 + nat
    return frames


def scf_options(**kw):
    o = ScfOptions()
    hfscf().hfscf_default_scf_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _callbacks(jk=None, grad2e=None, nbf=0, natom=0):
    """Wrap Python callables as an hfscf_callbacks struct (keeps references alive on the struct)."""

    def _jk(user, D, J, K):
        d = np.ctypeslib.as_array(D, shape=(nbf, nbf))
        j, k = jk(d.copy())
        np.ctypeslib.as_array(J, shape=(nbf, nbf))[:] = j
        np.ctypeslib.as_array(K, shape=(nbf, nbf))[:] = k
        return 0

    def _g(user, D, G):
        d = np.ctypeslib.as_array(D, shape=(nbf, nbf))
        np.ctypeslib.as_array(G, shape=(natom, 3))[:] = grad2e(d.copy())
        return 0

    cb = Callbacks()
    cb._keep = (JKFUNC(_jk), GRADFUNC(_g))
    cb.build_jk, cb.grad2e = cb._keep
    return cb


def rhf(mol, jk=None, grad2e=None, **opts):
    """RHF SCF (CORE guess). With jk=None the GPU backend builds every Fock matrix; a Python
    callable jk(D) -> (J, K) plugs into the same seam (used by tests to drive the host loop)."""
    lib = hfscf()
    o = scf_options(**opts)
    cb = None
    if jk is not None:
        o.backend = 1
        cb = _callbacks(jk, grad2e, mol.nbf, mol.natom)
    n = mol.nbf
    D, F = np.zeros((n, n)), np.zeros((n, n))
    E = np.zeros(max(o.max_iter, 1) + 1)
    res = ScfResult()
    rc = lib.hfscf_rhf_run(mol._h, C.byref(o), C.byref(cb) if cb else None, C.byref(res), _ptr(D), _ptr(F), _ptr(E))
    if rc not in (0, -2):
        raise RuntimeError(lib.hfscf_last_error().decode())
    return dict(energy=res.energy, energy_final=res.energy_final, energy_elec=res.energy_elec, energy_nuc=res.energy_nuc, iterations=res.iterations,
                converged=bool(res.converged), fock_seconds=res.fock_seconds, D=D, F=F,
                energies=E[:res.iterations + 1].copy())


def rhf_gradient(mol, D, F, jk=None, grad2e=None, **opts):
    """RHF analytic gradient (natom x 3, Eh/bohr)."""
    lib = hfscf()
    o = scf_options(**opts)
    cb = None
    if grad2e is not None:
        o.backend = 1
        cb = _callbacks(jk, grad2e, mol.nbf, mol.natom)
    g = np.zeros((mol.natom, 3))
    D = np.ascontiguousarray(D)
    F = np.ascontiguousarray(F)
    if lib.hfscf_rhf_gradient(mol._h, C.byref(o), C.byref(cb) if cb else None, _ptr(D), _ptr(F), _ptr(g)):
        raise RuntimeError(lib.hfscf_last_error().decode())
    return g


def batch_rhf(mols, gradient=False, **opts):
    """Batched RHF of many molecules (lockstep SCF, one batched GPU Fock build per iteration).
    Returns (list of per-molecule result dicts with D, F and optionally the gradient, stats)."""
    lib = hfscf()
    o = scf_options(**opts)
    n = len(mols)
    arr = (C.c_void_p * n)(*[m._h for m in mols])
    res = (ScfResult * n)()
    sizes = [m.nbf * m.nbf for m in mols]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    D = np.zeros(int(offs[-1]))
    F = np.zeros(int(offs[-1]))
    goffs = np.concatenate([[0], np.cumsum([3 * m.natom for m in mols])]).astype(np.int64)
    G = np.zeros(int(goffs[-1])) if gradient else None
    st = BatchStats()
    rc = lib.hfscf_batch_rhf(n, arr, C.byref(o), res, _ptr(D), _ptr(F), _ptr(G) if gradient else None, C.byref(st))
    if rc not in (0, -2):
        raise RuntimeError(lib.hfscf_last_error().decode())
    out = []
    for k, m in enumerate(mols):
        r = res[k]
        d = dict(energy=r.energy, energy_final=r.energy_final, iterations=r.iterations, converged=bool(r.converged),
                 D=D[offs[k]:offs[k + 1]].reshape(m.nbf, m.nbf), F=F[offs[k]:offs[k + 1]].reshape(m.nbf, m.nbf))
        if gradient:
            d["gradient"] = G[goffs[k]:goffs[k + 1]].reshape(m.natom, 3)
        out.append(d)
    return out, {k: getattr(st, k) for k, _ in BatchStats._fields_}


class GpuContext:
    """The GPU two-electron engine for one molecule, or for a batch of molecules when given a
    list (hfgpu C ABI; batched matrices are concatenated per molecule)."""

    def __init__(self, mol, precision=FP64, integral_tol=1e-14, schwarz=True, device=0, density_tol=0.0,
                 graph=False, work_split=0, coop_min_l=0, side_streams=0, group_shells=0):
        lib = hfgpu()
        opt = HfgOptions()
        lib.hfg_default_options(C.byref(opt))
        opt.precision, opt.integral_tol, opt.schwarz = precision, integral_tol, int(schwarz)
        opt.density_tol = density_tol
        opt.graph = int(graph)
        opt.work_split = int(work_split)
        opt.coop_min_l = int(coop_min_l)
        opt.side_streams = int(side_streams)
        opt.group_shells = int(group_shells)
        h = C.c_void_p()
        rc = lib.hfg_create(C.byref(h), device, C.byref(opt))
        if rc:
            raise RuntimeError(f"hfg_create failed with status {rc} (no usable GPU?)")
        self._h = h
        if isinstance(mol, (list, tuple)):
            self.mols = list(mol)
            self.mol = None
            for m in self.mols:
                sh = m.shells()
                x = np.ascontiguousarray(m.geometry()[1], dtype=np.float64).reshape(-1)
                rc = lib.hfg_batch_add_molecule(h, m.nshell, sh["L"].ctypes.data_as(_ip),
                                                sh["atom"].ctypes.data_as(_ip), sh["nprim"].ctypes.data_as(_ip),
                                                _ptr(sh["alpha"]), _ptr(sh["coef"]), int(m.pure),
                                                sh["cart_off"].ctypes.data_as(_ip),
                                                sh["pure_off"].ctypes.data_as(_ip), m.natom, _ptr(x))
                if rc < 0:
                    self._check(rc)
            self._check(lib.hfg_batch_commit(h))
            self.nbf = None
            self.sizes = [m.nbf * m.nbf for m in self.mols]
            self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
            self.natom = sum(m.natom for m in self.mols)
            return
        self.mols = [mol]
        self.mol = mol
        self.nbf = mol.nbf
        self.natom = mol.natom
        sh = mol.shells()
        self._check(lib.hfg_set_basis(h, mol.nshell, sh["L"].ctypes.data_as(_ip), sh["atom"].ctypes.data_as(_ip),
                                      sh["nprim"].ctypes.data_as(_ip), _ptr(sh["alpha"]), _ptr(sh["coef"]),
                                      int(mol.pure), sh["cart_off"].ctypes.data_as(_ip),
                                      sh["pure_off"].ctypes.data_as(_ip)))
        self.set_geometry(mol.geometry()[1])

    def _check(self, rc):
        if rc:
            raise RuntimeError(f"hfgpu status {rc}: {hfgpu().hfg_last_error(self._h).decode()}")

    def set_geometry(self, xyz_bohr):
        Z, _ = self.mol.geometry()
        x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1)
        self._check(hfgpu().hfg_set_geometry(self._h, len(Z), _ptr(x), Z.ctypes.data_as(_ip)))

    def _like(self, D):
        return np.zeros((self.nbf, self.nbf)) if self.nbf is not None else np.zeros(int(self.offsets[-1]))

    def build_jk(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64)
        J, K = self._like(D), self._like(D)
        self._check(hfgpu().hfg_build_jk(self._h, _ptr(D), _ptr(J), _ptr(K), 0))
        return J, K

    def fock(self, H, D, incremental=False):
        F = self._like(D)
        H = np.ascontiguousarray(H, dtype=np.float64)
        D = np.ascontiguousarray(D, dtype=np.float64)
        self._check(hfgpu().hfg_fock(self._h, _ptr(H), _ptr(D), int(incremental), _ptr(F)))
        return F

    def split(self, flat):
        """Per-molecule matrices of a concatenated (batch) array."""
        return [flat[self.offsets[k]:self.offsets[k + 1]].reshape(m.nbf, m.nbf) for k, m in enumerate(self.mols)]

    def class_stats(self):
        """Per class of the last direct J/K build: quartets, model FLOPs, kernel ms."""
        lib = hfgpu()
        n = lib.hfg_class_stats(self._h, 0, None, None, None, None)
        cls = np.zeros(n, dtype=np.int32)
        q = np.zeros(n, dtype=np.int64)
        fl, ms = np.zeros(n), np.zeros(n)
        self._check(min(0, lib.hfg_class_stats(self._h, n, cls.ctypes.data_as(_ip),
                                               q.ctypes.data_as(C.POINTER(C.c_longlong)), _ptr(fl), _ptr(ms))))
        return [dict(cls=f"{c:04d}", quartets=int(a), model_flops=float(b), kernel_ms=float(d))
                for c, a, b, d in zip(cls, q, fl, ms)]

    def autotune(self, what=1):
        self._check(hfgpu().hfg_autotune(self._h, int(what)))

    def launch_shapes(self):
        lib = hfgpu()
        n = lib.hfg_launch_shapes(self._h, 0, None, None, None, None, None)
        a = [np.zeros(n, dtype=np.int32) for _ in range(5)]
        lib.hfg_launch_shapes(self._h, n, *[x.ctypes.data_as(_ip) for x in a])
        return [dict(cls=f"{c:04d}", jk=(int(g), int(b)), grad=(int(gg), int(gb))) for c, g, b, gg, gb in zip(*a)]

    def set_precision(self, precision):
        self._check(hfgpu().hfg_set_precision(self._h, int(precision)))

    def set_density_tol(self, density_tol):
        self._check(hfgpu().hfg_set_density_tol(self._h, float(density_tol)))

    def energy(self, D):
        e = C.c_double()
        D = np.ascontiguousarray(D, dtype=np.float64)
        self._check(hfgpu().hfg_energy(self._h, _ptr(D), C.byref(e)))
        return e.value

    def energy_backward(self, D):
        g = np.zeros((self.natom, 3))
        D = np.ascontiguousarray(D, dtype=np.float64)
        self._check(hfgpu().hfg_energy_backward(self._h, _ptr(D), _ptr(g)))
        return g

    def oneint(self):
        """S, T, V of the molecule from the GPU one-electron kernels (single-molecule contexts)."""
        n = self.nbf or 1  # a batch context is refused by the library
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        self._check(hfgpu().hfg_oneint(self._h, _ptr(S), _ptr(T), _ptr(V)))
        return S, T, V

    def oneint_gradient(self, D, Q):
        """2 D.d(T+V) - 2 Q.dS (natom x 3) from the GPU one-electron kernels."""
        g = np.zeros((self.natom, 3))
        D = np.ascontiguousarray(D, dtype=np.float64)
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        self._check(hfgpu().hfg_oneint_gradient(self._h, _ptr(D), _ptr(Q), _ptr(g)))
        return g

    def eri_vector(self):
        """The packed (ij|kl) vector at hfscf's canonical index (single-molecule contexts)."""
        npair = self.nbf * (self.nbf + 1) // 2
        v = np.zeros(npair * (npair + 1) // 2)
        self._check(hfgpu().hfg_eri_vector(self._h, _ptr(v), len(v)))
        return v

    def hessian_2e(self, D):
        """Two-electron Hessian at the fixed density D: (HC, HX), each (3 natom) x (3 natom), as
        hfscf's m_hessianC and m_hessianEx (single-molecule contexts, shells up to d)."""
        n3 = 3 * self.natom
        hc, hx = np.zeros((n3, n3)), np.zeros((n3, n3))
        D = np.ascontiguousarray(D, dtype=np.float64)
        if D.shape != (self.nbf, self.nbf):
            raise ValueError(f"hessian_2e: D must be {self.nbf} x {self.nbf}")
        self._check(hfgpu().hfg_hessian_2e(self._h, _ptr(D), _ptr(hc), _ptr(hx)))
        return hc, hx

    def eri_block(self, s1, s2, s3, s4):
        sh = self.mol.shells()
        nf = [(2 * sh["L"][s] + 1) if self.mol.pure else (sh["L"][s] + 1) * (sh["L"][s] + 2) // 2
              for s in (s1, s2, s3, s4)]
        out = np.zeros(nf)
        self._check(hfgpu().hfg_eri_shell_quartet(self._h, s1, s2, s3, s4, _ptr(out)))
        return out

    def boys(self, m, T, closed_form=False):
        """F_0..F_m at the points T (len(T) x (m+1)) from the device Boys functions."""
        T = np.ascontiguousarray(T, dtype=np.float64).reshape(-1)
        F = np.zeros((len(T), m + 1))
        self._check(hfgpu().hfg_boys_probe(self._h, int(m), len(T), _ptr(T), _ptr(F), int(closed_form)))
        return F

    def schwarz(self):
        S = np.zeros((self.mol.ncart, self.mol.ncart))
        self._check(hfgpu().hfg_schwarz(self._h, _ptr(S)))
        return S

    def stats(self):
        st = HfgStats()
        hfgpu().hfg_last_stats(self._h, C.byref(st))
        return {k: getattr(st, k) for k, _ in HfgStats._fields_}

    def close(self):
        if getattr(self, "_h", None):
            hfgpu().hfg_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
