"""caps_amd — the MI355X force-and-step path of CAPS, reached from Python through the capsgpu C ABI.

The Python layer mirrors the reference's own interfaces for this path:
  * Evaluator(ff, options).compute(x, cell) -> (EnergyTerms, forces)   core/include/caps/field.hpp:327-352
  * run_dynamics(system, options) -> DynamicsReport                    core/include/caps/dynamics.hpp:180
with the same meanings, units and error behaviour (FieldError for force-field failures, ValueError for bad
arguments). Everything is computed by libcapsgpu.so (HIP kernels for gfx950); there is no CPU fallback: the module
raises when the library is missing or no GPU is visible.
"""
import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcapsgpu.so")

NTERMS = 15
TERM_NAMES = ("bond", "angle", "dihedral", "improper", "vdw", "coulomb", "virial", "wxx", "wyy", "wzz", "wxy", "wxz", "wyz",
              "tail_energy", "tail_virial")


class FieldError(RuntimeError):
    """caps::FieldError (core/include/caps/field.hpp:17-19)"""


class DeviceError(RuntimeError):
    """a HIP runtime failure or no GPU"""


class DynamicsCancelled(RuntimeError):
    """caps::DynamicsCancelled (core/include/caps/dynamics.hpp:168)"""


_lib = None


class FF(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_types", C.c_int32), ("type", C.c_void_p), ("charge", C.c_void_p), ("mass", C.c_void_p),
                ("eps", C.c_void_p), ("sigma", C.c_void_p), ("lj96", C.c_int32), ("excl_offset", C.c_void_p),
                ("excl_index", C.c_void_p), ("p14_offset", C.c_void_p), ("p14_index", C.c_void_p), ("coul14", C.c_double),
                ("n_bonds", C.c_int32), ("bond_atoms", C.c_void_p), ("bond_param", C.c_void_p), ("n_angles", C.c_int32),
                ("angle_atoms", C.c_void_p), ("angle_param", C.c_void_p), ("n_fene", C.c_int32), ("fene_atoms", C.c_void_p),
                ("fene_param", C.c_void_p)]


class EOpts(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("skin", C.c_double), ("coulomb", C.c_int32), ("electrostatics", C.c_int32),
                ("dsf_alpha", C.c_double), ("ewald_rtol", C.c_double), ("pme_spacing", C.c_double), ("pme_order", C.c_int32),
                ("tail", C.c_int32), ("parts", C.c_int32), ("no_reciprocal", C.c_int32)]


class DOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("steps", C.c_int64), ("temperature", C.c_double), ("thermostat", C.c_int32),
                ("tau_t", C.c_double), ("seed", C.c_uint64), ("thermo_every", C.c_int32), ("new_velocities", C.c_int32),
                ("respa", C.c_int32), ("step_offset", C.c_int64), ("learned_k", C.c_int32), ("learned_tol", C.c_double),
                ("learned_train", C.c_int32), ("barostat", C.c_int32), ("pressure", C.c_double), ("tau_p", C.c_double),
                ("compressibility", C.c_double), ("barostat_every", C.c_int32), ("temperature_end", C.c_double),
                ("checkpoint_every", C.c_int64), ("checkpoint", C.c_void_p), ("checkpoint_user", C.c_void_p)]


class Thermo(C.Structure):
    _fields_ = [("step", C.c_int64), ("time_ps", C.c_double), ("temperature", C.c_double), ("kinetic", C.c_double),
                ("potential", C.c_double), ("total", C.c_double), ("pressure", C.c_double), ("conserved", C.c_double),
                ("terms", C.c_double * NTERMS), ("learned", C.c_int32), ("learned_rmse", C.c_double),
                ("learned_rmse_stale", C.c_double), ("volume", C.c_double), ("density", C.c_double), ("p", C.c_double * 6)]


class DReport(C.Structure):
    _fields_ = [("steps", C.c_int64), ("seconds", C.c_double), ("ns_per_day", C.c_double), ("list_builds", C.c_int32),
                ("prunes", C.c_int32), ("learned_steps", C.c_int32), ("full_steps", C.c_int32), ("fallbacks", C.c_int32),
                ("judgments", C.c_int32)]


class BOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("steps", C.c_int64), ("tau_t", C.c_double), ("seed", C.c_uint64), ("new_velocities", C.c_int32),
                ("mode", C.c_int32), ("branches", C.c_int32), ("graph_steps", C.c_int32)]


class BReport(C.Structure):
    _fields_ = [("steps", C.c_int64), ("seconds", C.c_double), ("atom_steps_per_s", C.c_double), ("ns_per_day", C.c_double),
                ("mode", C.c_int32), ("branches", C.c_int32), ("graph_steps", C.c_int32), ("tune_seconds", C.c_double)]


class COpts(C.Structure):
    _fields_ = [("t_start", C.c_double), ("t_end", C.c_double), ("t_step", C.c_double), ("ps_per_step", C.c_double),
                ("equilibrate_ps", C.c_double), ("average_from", C.c_double), ("dt", C.c_double), ("tau_t", C.c_double),
                ("tau_p", C.c_double), ("pressure", C.c_double), ("compressibility", C.c_double), ("barostat", C.c_int32),
                ("barostat_every", C.c_int32), ("seed", C.c_uint64), ("new_velocities", C.c_int32), ("property", C.c_int32),
                ("fit", C.c_int32), ("glassy_max", C.c_double), ("rubbery_min", C.c_double)]


class CPoint(C.Structure):
    _fields_ = [("temperature", C.c_double), ("density", C.c_double), ("density_err", C.c_double), ("specific_volume", C.c_double),
                ("potential", C.c_double)]


class CFit(C.Structure):
    _fields_ = [("ok", C.c_int32), ("note", C.c_int32), ("tg", C.c_double), ("tg_err", C.c_double), ("slope_low", C.c_double),
                ("slope_high", C.c_double), ("value_at_tg", C.c_double), ("alpha_low", C.c_double), ("alpha_high", C.c_double),
                ("rss", C.c_double)]


class ROpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("ftol", C.c_double), ("etol", C.c_double), ("max_iterations", C.c_int32)]


class RStage(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("evaluations", C.c_int32), ("stopped_by", C.c_int32), ("energy", C.c_double),
                ("fmax", C.c_double)]


class EqStage(C.Structure):
    _fields_ = [("ensemble", C.c_int32), ("ps", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double), ("pressure", C.c_double)]


class EqOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("tau_t", C.c_double), ("tau_p", C.c_double), ("compressibility", C.c_double),
                ("thermostat", C.c_int32), ("barostat", C.c_int32), ("barostat_every", C.c_int32), ("seed", C.c_uint64),
                ("thermo_ps", C.c_double)]


class EqSummary(C.Structure):
    _fields_ = [("ps", C.c_double), ("temperature", C.c_double), ("pressure", C.c_double), ("density", C.c_double),
                ("potential", C.c_double), ("density_end", C.c_double)]


class ProtoParams(C.Structure):
    _fields_ = [("t_final", C.c_double), ("t_max", C.c_double), ("p_final", C.c_double), ("p_max", C.c_double),
                ("time_scale", C.c_double), ("cycles", C.c_int32), ("t_low", C.c_double), ("t_high", C.c_double),
                ("ramp_ps", C.c_double), ("hold_ps", C.c_double)]


class Sys(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_types", C.c_int32), ("cell", C.c_double * 12), ("x", C.POINTER(C.c_double)),
                ("type", C.POINTER(C.c_int32)), ("charge", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
                ("eps", C.POINTER(C.c_double)), ("sigma", C.POINTER(C.c_double)), ("n_bonds", C.c_int32),
                ("bond_atoms", C.POINTER(C.c_int32)), ("bond_param", C.POINTER(C.c_double)), ("n_angles", C.c_int32),
                ("angle_atoms", C.POINTER(C.c_int32)), ("angle_param", C.POINTER(C.c_double)), ("n_fene", C.c_int32),
                ("fene_atoms", C.POINTER(C.c_int32)), ("fene_param", C.POINTER(C.c_double)), ("excl_offset", C.POINTER(C.c_int32)),
                ("excl_index", C.POINTER(C.c_int32)), ("v", C.POINTER(C.c_double)), ("opaque", C.c_void_p)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int32, C.POINTER(Thermo), C.c_void_p)
CHECKPOINT_FN = C.CFUNCTYPE(C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64, C.c_void_p)
COOLING_FN = C.CFUNCTYPE(C.c_int32, C.POINTER(Thermo), C.c_int32, C.c_int32, C.c_void_p)


def lib():
    """the loaded libcapsgpu.so; raises when it is not built"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcapsgpu.so is not built: run `python -m caps_amd.build` (hipcc, gfx950)")
        # one HIP runtime per process: when PyTorch is installed its libamdhip64 is loaded first, so the library binds
        # to the same runtime as the process's tensors and streams rather than bringing a second copy
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        L.capsgpu_last_error.restype = C.c_char_p
        L.capsgpu_build_error.restype = C.c_char_p
        L.capsgpu_create.argtypes = [C.POINTER(FF), C.POINTER(EOpts), C.POINTER(C.c_void_p)]
        L.capsgpu_destroy.argtypes = [C.c_void_p]
        L.capsgpu_compute.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.capsgpu_compute_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.capsgpu_list_builds.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.capsgpu_set_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.capsgpu_get_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.capsgpu_get_forces.argtypes = [C.c_void_p, C.c_void_p]
        L.capsgpu_run_dynamics.argtypes = [C.c_void_p, C.POINTER(DOpts), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                           C.POINTER(C.c_int32), C.POINTER(DReport)]
        L.capsgpu_list_sizes.argtypes = [C.c_void_p, C.c_void_p]
        L.capsgpu_list_get.argtypes = [C.c_void_p] + [C.c_void_p] * 13
        L.capsgpu_list_pruned.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.capsgpu_stage_times.argtypes = [C.c_void_p, C.c_void_p]
        L.capsgpu_profile.argtypes = [C.c_void_p, C.c_int32]
        L.capsgpu_profile_read.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.capsgpu_profile_kernels.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
        L.capsgpu_profile_only.argtypes = [C.c_void_p, C.c_char_p]
        L.capsgpu_learned_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.capsgpu_learned_gradient.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                               C.POINTER(C.c_double)]
        L.capsgpu_learned_descriptors.argtypes = [C.c_void_p, C.c_void_p]
        L.capsgpu_tune.argtypes = [C.c_void_p, C.c_void_p]
        L.capsgpu_set_tiles.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.capsgpu_set_gather.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.capsgpu_get_gather.argtypes = [C.c_void_p, C.c_void_p]
        L.capsgpu_batch_create.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
        L.capsgpu_batch_run.argtypes = [C.c_void_p, C.POINTER(BOpts), C.c_void_p, C.c_void_p, C.POINTER(BReport)]
        L.capsgpu_batch_destroy.argtypes = [C.c_void_p]
        L.capsgpu_minimise.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(ROpts), C.c_void_p, C.POINTER(RStage)]
        L.capsgpu_equilibrate.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(EqOpts), C.c_void_p, C.c_void_p, C.c_void_p]
        L.capsgpu_protocol.argtypes = [C.c_char_p, C.POINTER(ProtoParams), C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        L.capsgpu_run_cooling.argtypes = [C.c_void_p, C.POINTER(COpts), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                          C.POINTER(C.c_int32), C.POINTER(CFit)]
        L.capsgpu_batch_cooling.argtypes = [C.c_void_p, C.POINTER(COpts), C.c_int32, C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                            C.c_void_p]
        L.capsgpu_tg_fit_points.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double, C.POINTER(CFit)]
        L.capsgpu_build_melt.argtypes = [C.c_int32, C.c_uint64, C.c_double, C.POINTER(C.POINTER(Sys))]
        L.capsgpu_build_kg.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.POINTER(C.POINTER(Sys))]
        L.capsgpu_system_free.argtypes = [C.POINTER(Sys)]
        L.capsgpu_read_lammps.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.POINTER(Sys))]
        L.capsgpu_write_lammps.argtypes = [C.c_char_p, C.POINTER(Sys), C.c_void_p, C.c_char_p]
        _lib = L
    return _lib


def _check(rc):
    if rc == 0:
        return
    msg = lib().capsgpu_last_error().decode()
    if rc == -2:
        raise FieldError(msg)
    if rc == -3:
        raise DeviceError(msg)
    if rc == -4:
        raise DynamicsCancelled(msg)
    raise ValueError(msg)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---- data model (the parts of caps::ForceField / caps::System this path reads) ----------------------------------------
@dataclass
class ForceField:
    """the flattened force field: per atom type, charge, mass; per type pair mixed ε and σ; bonded terms; exclusions"""
    type: np.ndarray
    charge: np.ndarray
    mass: np.ndarray
    eps: np.ndarray            # nt x nt
    sigma: np.ndarray          # nt x nt
    excl_offset: np.ndarray
    excl_index: np.ndarray
    bond_atoms: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bond_param: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    angle_atoms: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    angle_param: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    fene_atoms: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    fene_param: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    p14_offset: np.ndarray = None
    p14_index: np.ndarray = None
    coul14: float = 1.0 / 1.2
    lj96: bool = False
    lj_shift: bool = False     # ForceField::lj_shift: LJ shifted to zero at the cut-off, no tail correction
    cutoff: float = 0.0        # ForceField::cutoff: the model's own cut-off (0: the options')

    @property
    def n(self):
        return int(self.type.size)

    @property
    def ntypes(self):
        return int(self.eps.shape[0])


@dataclass
class System:
    x: np.ndarray              # (N, 3) Å
    cell: np.ndarray           # (12,) origin, a, b, c
    ff: ForceField
    v: np.ndarray = None       # (N, 3) Å/fs
    title: str = ""            # a data file's first line


@dataclass
class EnergyOptions:
    """caps::EnergyOptions (core/include/caps/field.hpp:274-308)"""
    cutoff: float = 10.0
    skin: float = 2.0
    coulomb: bool = True
    electrostatics: str = "dsf"   # "dsf" or "pme"
    dsf_alpha: float = 0.2
    ewald_rtol: float = 1e-5
    pme_spacing: float = 1.0
    pme_order: int = 5
    tail: bool = True
    parts: int = 3                # 1 bonded, 2 non-bonded (pairs, k-space, self, tail), 3 both; 4 PME's reciprocal part
                                  # with its self and background terms alone (the slow force of learned r-RESPA)
    no_reciprocal: bool = False   # checks only: the real-space part of PME alone


@dataclass
class EnergyTerms:
    """caps::EnergyTerms (core/include/caps/field.hpp:310-320)"""
    bond: float = 0.0
    angle: float = 0.0
    dihedral: float = 0.0
    improper: float = 0.0
    vdw: float = 0.0
    coulomb: float = 0.0
    virial: float = 0.0
    w: tuple = (0.0,) * 6
    tail_energy: float = 0.0
    tail_virial: float = 0.0

    def total(self):
        return self.bond + self.angle + self.dihedral + self.improper + self.vdw + self.coulomb

    @staticmethod
    def from_array(t):
        t = [float(v) for v in t]
        return EnergyTerms(t[0], t[1], t[2], t[3], t[4], t[5], t[6], tuple(t[7:13]), t[13], t[14])


class Evaluator:
    """caps::Evaluator on one GPU (core/include/caps/field.hpp:327-352): holds the device copy of the force field and
    of the pair list between calls; compute() returns the energy terms and the forces."""

    def __init__(self, ff, options=None):
        self.ff = ff
        o = options or EnergyOptions()
        self.options = o
        L = lib()
        f = FF()
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        f.n_atoms, f.n_types = ff.n, ff.ntypes
        f.type, f.charge, f.mass = arr(ff.type, np.int32), arr(ff.charge, np.float64), arr(ff.mass, np.float64)
        f.eps, f.sigma = arr(ff.eps, np.float64), arr(ff.sigma, np.float64)
        f.lj96 = int(ff.lj96)
        f.excl_offset, f.excl_index = arr(ff.excl_offset, np.int32), arr(ff.excl_index if ff.excl_index.size else [0], np.int32)
        if ff.p14_offset is not None:
            f.p14_offset, f.p14_index = arr(ff.p14_offset, np.int32), arr(ff.p14_index if ff.p14_index.size else [0], np.int32)
        f.coul14 = ff.coul14
        f.n_bonds = len(ff.bond_atoms)
        f.bond_atoms, f.bond_param = arr(ff.bond_atoms, np.int32), arr(ff.bond_param, np.float64)
        f.n_angles = len(ff.angle_atoms)
        f.angle_atoms, f.angle_param = arr(ff.angle_atoms, np.int32), arr(ff.angle_param, np.float64)
        f.n_fene = len(ff.fene_atoms)
        f.fene_atoms, f.fene_param = arr(ff.fene_atoms, np.int32), arr(ff.fene_param, np.float64)
        e = EOpts()
        e.cutoff = ff.cutoff if ff.cutoff > 0 else o.cutoff
        e.skin, e.coulomb = o.skin, int(o.coulomb)
        e.electrostatics = 1 if o.electrostatics.lower() == "pme" else 0
        e.dsf_alpha, e.ewald_rtol, e.pme_spacing, e.pme_order = o.dsf_alpha, o.ewald_rtol, o.pme_spacing, o.pme_order
        e.tail = int(o.tail and not ff.lj_shift)   # field.cpp:442: lj_shift turns the tail correction off
        e.parts = o.parts
        e.no_reciprocal = int(o.no_reciprocal)
        self.cutoff = e.cutoff
        self._h = C.c_void_p()
        _check(L.capsgpu_create(C.byref(f), C.byref(e), C.byref(self._h)))
        self._energies = True

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().capsgpu_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def energies(self, on):
        """Evaluator::energies(bool): skip pair energies and the virial on steps that report none"""
        self._energies = bool(on)

    def compute(self, x, cell):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        cell = np.ascontiguousarray(cell, np.float64)
        f = np.zeros_like(x)
        t = np.zeros(NTERMS)
        _check(lib().capsgpu_compute(self._h, _p(x), _p(cell), _p(f), _p(t), int(self._energies)))
        return EnergyTerms.from_array(t), f.reshape(-1, 3)

    def list_builds(self):
        b = C.c_int32()
        _check(lib().capsgpu_list_builds(self._h, C.byref(b)))
        return b.value

    def stage_times(self):
        t = np.zeros(6)
        _check(lib().capsgpu_stage_times(self._h, _p(t)))
        return dict(zip(("list_build", "prune", "pairs", "pme", "bonded", "rest"), t))

    def profile(self, on=True):
        """start (zeroing the sums) or stop the stage profiling of every later evaluation, dynamics included"""
        _check(lib().capsgpu_profile(self._h, int(on)))

    def profile_only(self, kernel=None):
        """time only the launches of one kernel (None: every kernel) and no stages"""
        _check(lib().capsgpu_profile_only(self._h, kernel.encode() if kernel else None))

    def tune(self):
        """picks this context's mesh-transform tile shape by timing (no reference counterpart); returns (line pairs along
        a, lines along b and c) per workgroup. Needs PME and a set state (set_state or a compute first)."""
        t = np.zeros(2, np.int32)
        _check(lib().capsgpu_tune(self._h, _p(t)))
        return int(t[0]), int(t[1])

    def gather_shape(self):
        """the force gather's workgroups in use: 0 one per 4x4x4 tile, 1 one per run of 64 sorted atoms (picked by tune())"""
        g = np.zeros(1, np.int32)
        _check(lib().capsgpu_get_gather(self._h, _p(g)))
        return int(g[0])

    def set_gather(self, shape):
        """sets the force gather's workgroups (0 per tile, 1 per run of 64 sorted atoms, one atom per lane, where K1 is a
        multiple of 4); the forces are the same either way"""
        g = np.zeros(1, np.int32)
        _check(lib().capsgpu_set_gather(self._h, int(shape), _p(g)))
        return int(g[0])

    def set_tiles(self, lines_x, lines_yz):
        """sets the mesh-transform tile shape (powers of two, 1 to 16 lines)"""
        t = np.zeros(2, np.int32)
        _check(lib().capsgpu_set_tiles(self._h, int(lines_x), int(lines_yz), _p(t)))
        return int(t[0]), int(t[1])

    def profile_read(self):
        """{stage: (summed device ms, evaluations)} since profiling started"""
        ms, n = np.zeros(6), np.zeros(6, np.int64)
        _check(lib().capsgpu_profile_read(self._h, _p(ms), _p(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(("list_build", "prune", "pairs", "pme", "bonded", "rest"))}

    # learned r-RESPA inspection (parity checks)
    def learned_forward(self, params, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), 3), np.float32)
        _check(lib().capsgpu_learned_forward(self._h, _p(np.ascontiguousarray(params, np.float32)), _p(x), len(x), _p(y)))
        return y

    def learned_gradient(self, params, x, t):
        x, t = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(t, np.float32)
        g = np.zeros(len(params), np.float32)
        loss = C.c_double()
        _check(lib().capsgpu_learned_gradient(self._h, _p(np.ascontiguousarray(params, np.float32)), _p(x), _p(t), len(x), _p(g),
                                              C.byref(loss)))
        return loss.value, g

    def learned_descriptors(self):
        d = np.zeros((self.ff.n, 96), np.float32)
        _check(lib().capsgpu_learned_descriptors(self._h, _p(d)))
        return d

    def profile_kernels(self):
        """{kernel: (summed device ms, launches)} since profiling started"""
        buf = C.create_string_buffer(1 << 16)
        _check(lib().capsgpu_profile_kernels(self._h, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n = line.split()
            out[name] = (float(ms), int(n))
        return out

    # inspection of the cluster list (parity checks)
    def list_arrays(self):
        L = lib()
        s = np.zeros(8, np.int64)
        _check(L.capsgpu_list_sizes(self._h, _p(s)))
        n, ncl, ns, ne, nc = (int(v) for v in s[:5])
        out = dict(perm=np.zeros(ns, np.uint32), iperm=np.zeros(n, np.uint32), first=np.zeros(ncl + 1, np.int64),
                   cj=np.zeros(ne, np.uint32), code=np.zeros(ne, np.uint16), mask=np.zeros(ne, np.uint16), ex=np.zeros(ne, np.uint16),
                   ex14=np.zeros(ne, np.uint16), org=np.zeros(3 * ncl), off=np.zeros(3 * ns), ioff=np.zeros(3 * ns, np.int32),
                   qint=np.zeros(3 * nc, np.int32), qvec=np.zeros(3 * nc))
        keys = ("perm", "iperm", "first", "cj", "code", "mask", "ex", "ex14", "org", "off", "ioff", "qint", "qvec")
        _check(L.capsgpu_list_get(self._h, *[_p(out[k]) for k in keys]))
        out["nu"], out["nv"], out["npruned"] = int(s[5]), int(s[6]), int(s[7])
        return out

    def list_pruned(self):
        L = lib()
        s = np.zeros(8, np.int64)
        _check(L.capsgpu_list_sizes(self._h, _p(s)))
        ns, ne = int(s[2]), int(s[3])
        X, Y, Z = (np.zeros(ns, np.float32) for _ in range(3))
        pm, pe, p14 = (np.zeros(ne, np.uint16) for _ in range(3))
        _check(L.capsgpu_list_pruned(self._h, _p(X), _p(Y), _p(Z), _p(pm), _p(pe), _p(p14)))
        return X, Y, Z, pm, pe, p14

    # dynamics state on the device
    def set_state(self, x, v=None, cell=None):
        x = np.ascontiguousarray(x, np.float64).reshape(-1)
        v = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        cell = None if cell is None else np.ascontiguousarray(cell, np.float64)
        _check(lib().capsgpu_set_state(self._h, _p(x), _p(v), _p(cell)))

    def get_state(self):
        n = self.ff.n
        x, v, cell = np.zeros(3 * n), np.zeros(3 * n), np.zeros(12)
        _check(lib().capsgpu_get_state(self._h, _p(x), _p(v), _p(cell)))
        return x.reshape(-1, 3), v.reshape(-1, 3), cell


@dataclass
class DynamicsOptions:
    """caps::DynamicsOptions (core/include/caps/dynamics.hpp:60-154), the fields this path reads"""
    dt: float = 1.0
    steps: int = 10000
    temperature: float = 300.0
    thermostat: str = "bussi"        # none, bussi, langevin
    tau_t: float = 100.0
    seed: int = 1
    thermo_every: int = 100
    new_velocities: bool = False
    respa: int = 1
    step_offset: int = 0
    energy: EnergyOptions = field(default_factory=EnergyOptions)
    progress: object = None          # callable(row) -> bool
    learned_k: int = 0                # learned r-RESPA: a full PME evaluation every k steps (0: off)
    learned_tol: float = 1e-3         # relative force RMSE above which the next interval runs plain PME
    learned_train: int = 2000         # PME frames the network trains on before it predicts
    barostat: str = "none"            # none, berendsen, c-rescale (isotropic; dynamics.cpp:701-723)
    pressure: float = 1.0             # atm
    tau_p: float = 1000.0             # fs
    compressibility: float = 4.5e-5   # 1/atm
    barostat_every: int = 10          # steps between cell updates
    temperature_end: float = -1.0     # K; > 0 ramps the thermostat target over the run
    checkpoint_every: int = 0         # steps between checkpoints (0: none; dynamics.hpp:143-146)
    checkpoint: object = None         # a path (a LAMMPS data file with velocities, written whole) or callable(x, v, cell, step)


@dataclass
class DynamicsReport:
    thermo: list
    steps: int
    seconds: float
    ns_per_day: float
    list_builds: int
    prunes: int
    learned_steps: int = 0
    full_steps: int = 0
    fallbacks: int = 0               # learned r-RESPA: judgments that missed the tolerance
    judgments: int = 0               # learned r-RESPA: full steps that judged a k-step prediction


_THERMOSTATS = {"none": 0, "nve": 0, "bussi": 1, "csvr": 1, "v-rescale": 1, "langevin": 2, "baoab": 2}
_BAROSTATS = {"none": 0, "berendsen": 1, "c-rescale": 2, "crescale": 2}


def run_dynamics(system, o, evaluator=None):
    """caps::run_dynamics (core/src/dynamics.cpp:71-815) on the GPU: updates system.x, system.v in place and returns a
    DynamicsReport with the thermo rows (dicts)."""
    if o.thermostat.lower() not in _THERMOSTATS:
        raise ValueError("unknown thermostat '%s' (none, bussi, langevin)" % o.thermostat)
    ev = evaluator or Evaluator(system.ff, o.energy)
    has_v = system.v is not None and not o.new_velocities
    ev.set_state(system.x, system.v if has_v else None, system.cell)
    d = DOpts()
    d.dt, d.steps, d.temperature = o.dt, o.steps, o.temperature
    d.thermostat, d.tau_t, d.seed = _THERMOSTATS[o.thermostat.lower()], o.tau_t, o.seed
    d.thermo_every, d.new_velocities, d.respa, d.step_offset = o.thermo_every, int(not has_v), o.respa, o.step_offset
    d.learned_k, d.learned_tol, d.learned_train = o.learned_k, o.learned_tol, o.learned_train
    if o.barostat.lower() not in _BAROSTATS:
        raise ValueError("unknown barostat '%s' (none, berendsen, c-rescale)" % o.barostat)
    d.barostat, d.pressure, d.tau_p = _BAROSTATS[o.barostat.lower()], o.pressure, o.tau_p
    d.compressibility, d.barostat_every = o.compressibility, o.barostat_every
    d.temperature_end = o.temperature_end
    cap = int(o.steps // max(1, o.thermo_every) + 2)
    rows = (Thermo * cap)()
    nr = C.c_int32()
    rep = DReport()
    failure = []
    cpf = None
    if o.checkpoint_every > 0 and o.checkpoint is not None:
        n = system.ff.n
        total = o.step_offset + o.steps

        def _cp(xp, vp, cp, step, user):
            try:
                x = np.ctypeslib.as_array(xp, shape=(3 * n,)).reshape(-1, 3).copy()
                v = np.ctypeslib.as_array(vp, shape=(3 * n,)).reshape(-1, 3).copy()
                cell = np.ctypeslib.as_array(cp, shape=(12,)).copy()
                if callable(o.checkpoint):
                    o.checkpoint(x, v, cell, int(step))
                else:
                    write_checkpoint(o.checkpoint, System(x=x, cell=cell, ff=system.ff, v=v), int(step), total)
                return 0
            except Exception as e:   # noqa: BLE001 - handed back after the run
                failure.append(e)
                return 1
        cpf = CHECKPOINT_FN(_cp)
        d.checkpoint_every, d.checkpoint = o.checkpoint_every, C.cast(cpf, C.c_void_p)
    cb = None
    if o.progress is not None:
        def _cb(row, user):
            return 1 if o.progress(_row(row.contents)) else 0
        cb = PROGRESS_FN(_cb)
    rc = lib().capsgpu_run_dynamics(ev._h, C.byref(d), C.cast(cb, C.c_void_p) if cb else None, None, C.cast(rows, C.c_void_p), cap,
                                    C.byref(nr), C.byref(rep))
    if failure:
        raise failure[0]
    _check(rc)
    x, v, cell = ev.get_state()
    system.x, system.v, system.cell = x, v, cell
    return DynamicsReport([_row(rows[k]) for k in range(min(nr.value, cap))], rep.steps, rep.seconds, rep.ns_per_day,
                          rep.list_builds, rep.prunes, rep.learned_steps, rep.full_steps, rep.fallbacks, rep.judgments)


@dataclass
class BatchOptions:
    """Langevin BAOAB steps of many replicas together (no reference counterpart: the reference runs a cooling scan's
    replicas one after another through run_dynamics, core/src/mechanics.cpp:1381-1470). mode: "tuned" (the launch shape
    timed on the batch), "graph" (branches streams, graph_steps steps per captured graph), "direct" (no graph) or
    "grouped" (each kernel of the step launched once for all replicas, graph_steps steps per graph)."""
    dt: float = 1.0
    steps: int = 1000
    tau_t: float = 100.0
    seed: int = 1                     # replica r draws from seed + r
    new_velocities: bool = True
    mode: str = "tuned"
    branches: int = 1
    graph_steps: int = 1


@dataclass
class BatchReport:
    steps: int
    seconds: float
    atom_steps_per_s: float
    ns_per_day: float
    mode: str
    branches: int
    graph_steps: int
    tune_seconds: float
    temperatures: np.ndarray          # the kinetic temperature of each replica at the end


_BATCH_MODES = {"tuned": 0, "graph": 1, "direct": 2, "grouped": 3}


class Batch:
    """Replicas (one Evaluator each: its own force field, cell, PME grid) stepped together. Each replica ends as
    run_dynamics(system, DynamicsOptions(thermostat="langevin", seed=seed + r, ...)) of that replica alone would leave it,
    bit for bit. The evaluators' options must leave the pair path out (parts 1, 4 or 5)."""

    def __init__(self, evaluators):
        self.evaluators = list(evaluators)
        n = len(self.evaluators)
        hs = (C.c_void_p * n)(*[e._h.value for e in self.evaluators])
        self._h = C.c_void_p()
        _check(lib().capsgpu_batch_create(C.cast(hs, C.c_void_p), n, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().capsgpu_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, systems, temperatures, o=None):
        """steps every system from its state at its temperature (K); updates system.x and system.v in place"""
        o = o or BatchOptions()
        if o.mode not in _BATCH_MODES:
            raise ValueError("unknown batch mode '%s' (tuned, graph, direct)" % o.mode)
        if len(systems) != len(self.evaluators) or len(temperatures) != len(self.evaluators):
            raise ValueError("one system and one temperature per replica")
        if not o.new_velocities and any(s.v is None for s in systems):
            raise ValueError("every replica needs velocities unless new_velocities is set")
        for ev, s in zip(self.evaluators, systems):
            ev.set_state(s.x, None if o.new_velocities else s.v, s.cell)
        b = BOpts()
        b.dt, b.steps, b.tau_t, b.seed, b.new_velocities = o.dt, o.steps, o.tau_t, o.seed, int(o.new_velocities)
        b.mode, b.branches, b.graph_steps = _BATCH_MODES[o.mode], o.branches, o.graph_steps
        T = np.ascontiguousarray(temperatures, dtype=np.float64)
        Tout = np.zeros(len(self.evaluators))
        r = BReport()
        _check(lib().capsgpu_batch_run(self._h, C.byref(b), T.ctypes.data_as(C.c_void_p), Tout.ctypes.data_as(C.c_void_p), C.byref(r)))
        for ev, s in zip(self.evaluators, systems):
            x, v, _ = ev.get_state()
            s.x, s.v = x, v
        mode = {1: "graph", 2: "direct", 3: "grouped"}.get(r.mode, "tuned")
        return BatchReport(r.steps, r.seconds, r.atom_steps_per_s, r.ns_per_day, mode, r.branches, r.graph_steps, r.tune_seconds, Tout)


def _row(r):
    return dict(step=r.step, time_ps=r.time_ps, temperature=r.temperature, kinetic=r.kinetic, potential=r.potential, total=r.total,
                pressure=r.pressure, conserved=r.conserved, terms=EnergyTerms.from_array(r.terms), learned=r.learned,
                learned_rmse=r.learned_rmse, learned_rmse_stale=r.learned_rmse_stale, volume=r.volume, density=r.density,
                p=tuple(r.p))


# ---- synthetic systems ------------------------------------------------------------------------------------------------
def _system_from(ptr):
    s = ptr.contents
    n, nt = s.n_atoms, s.n_types

    def a(p, m, dt):
        return np.ctypeslib.as_array(p, shape=(m,)).astype(dt).copy() if m else np.zeros(0, dt)

    x = a(s.x, 3 * n, np.float64).reshape(-1, 3)
    off = a(s.excl_offset, n + 1, np.int32)
    ff = ForceField(type=a(s.type, n, np.int32), charge=a(s.charge, n, np.float64), mass=a(s.mass, n, np.float64),
                    eps=a(s.eps, nt * nt, np.float64).reshape(nt, nt), sigma=a(s.sigma, nt * nt, np.float64).reshape(nt, nt),
                    excl_offset=off, excl_index=a(s.excl_index, int(off[-1]), np.int32),
                    bond_atoms=a(s.bond_atoms, 2 * s.n_bonds, np.int32).reshape(-1, 2),
                    bond_param=a(s.bond_param, 2 * s.n_bonds, np.float64).reshape(-1, 2),
                    angle_atoms=a(s.angle_atoms, 3 * s.n_angles, np.int32).reshape(-1, 3),
                    angle_param=a(s.angle_param, 2 * s.n_angles, np.float64).reshape(-1, 2),
                    fene_atoms=a(s.fene_atoms, 2 * s.n_fene, np.int32).reshape(-1, 2),
                    fene_param=a(s.fene_param, 4 * s.n_fene, np.float64).reshape(-1, 4))
    cell = np.array(list(s.cell), np.float64)
    v = a(s.v, 3 * n, np.float64).reshape(-1, 3) if s.v else None
    lib().capsgpu_system_free(ptr)
    return System(x=x, cell=cell, ff=ff, v=v)


def build_melt(chains, seed=1, density=1.0):
    """an all-atom polyethylene-like melt of 128-atom chains (the 32,768- and 262,144-atom cases)"""
    p = C.POINTER(Sys)()
    if lib().capsgpu_build_melt(chains, seed, density, C.byref(p)) != 0:
        raise ValueError(lib().capsgpu_build_error().decode())
    return _system_from(p)


def kremer_grest(chains=200, beads=50, density=0.85, k_theta=0.0, bond=0.97, seed=1):
    """kremer_grest() + kremer_grest_forcefield() (core/src/kremer_grest.cpp:15-124): FENE + WCA, r_c = 2^(1/6)"""
    p = C.POINTER(Sys)()
    if lib().capsgpu_build_kg(chains, beads, density, k_theta, bond, seed, C.byref(p)) != 0:
        raise ValueError(lib().capsgpu_build_error().decode())
    s = _system_from(p)
    s.ff.lj_shift = True
    s.ff.cutoff = 2.0 ** (1.0 / 6.0)
    s.ff.coul14 = 1.0
    return s


# ---- LAMMPS data files and checkpoints --------------------------------------------------------------------------------
def _settings_path(path):
    """the run settings beside a data file: X.restart.data -> X.restart.settings.json (cli/main.cpp:2229-2231)"""
    return os.path.splitext(path)[0] + ".settings.json"


def _excludes_13(ff):
    """whether the force field excludes the end atoms of its angles (the melt's 1-3 exclusions; Kremer-Grest has none)"""
    if len(ff.angle_atoms) == 0:
        return False
    i, k = int(ff.angle_atoms[0][0]), int(ff.angle_atoms[0][2])
    return k in ff.excl_index[ff.excl_offset[i]:ff.excl_offset[i + 1]]


def read_lammps(path, exclude13=None):
    """read_lammps_data (core/src/io_lammps.cpp:58-297) for atom_style full, units real, with this path's force field
    from the coefficient sections. The run settings a data file cannot hold (LJ shifted at its own cut-off, 1-4 Coulomb
    scale, 1-3 exclusions) are taken from the settings file a checkpoint writes beside it, when there is one."""
    import json
    settings = {}
    sp = _settings_path(path)
    if os.path.exists(sp):
        with open(sp) as f:
            settings = json.load(f)
    if exclude13 is None:
        exclude13 = settings.get("exclude13", None)
    p = C.POINTER(Sys)()
    _check(lib().capsgpu_read_lammps(os.fsencode(path), -1 if exclude13 is None else int(bool(exclude13)), C.byref(p)))
    s = _system_from(p)
    with open(path, errors="replace") as f:
        s.title = f.readline().strip()
    for k in ("lj_shift", "cutoff", "coul14", "lj96"):
        if k in settings:
            setattr(s.ff, k, type(getattr(s.ff, k))(settings[k]))
    return s


def write_lammps(path, system, title="", velocities=True):
    """write_lammps_data (core/src/io_lammps.cpp:505-583) with the coefficient sections of this path's force field and
    17 significant digits, so the file reads back to the same numbers; written to path.tmp, then renamed."""
    ff = system.ff
    keep = []

    def arr(a, dt, ptr):
        a = np.ascontiguousarray(a, dtype=dt).reshape(-1)
        if a.size == 0:
            a = np.zeros(1, dt)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(ptr))

    S = Sys()
    S.n_atoms, S.n_types = ff.n, ff.ntypes
    S.cell = (C.c_double * 12)(*np.asarray(system.cell, np.float64).reshape(12))
    S.x = arr(system.x, np.float64, C.c_double)
    S.type, S.charge, S.mass = arr(ff.type, np.int32, C.c_int32), arr(ff.charge, np.float64, C.c_double), arr(ff.mass, np.float64, C.c_double)
    S.eps, S.sigma = arr(ff.eps, np.float64, C.c_double), arr(ff.sigma, np.float64, C.c_double)
    S.n_bonds, S.bond_atoms, S.bond_param = len(ff.bond_atoms), arr(ff.bond_atoms, np.int32, C.c_int32), arr(ff.bond_param, np.float64, C.c_double)
    S.n_angles, S.angle_atoms = len(ff.angle_atoms), arr(ff.angle_atoms, np.int32, C.c_int32)
    S.angle_param = arr(ff.angle_param, np.float64, C.c_double)
    S.n_fene, S.fene_atoms, S.fene_param = len(ff.fene_atoms), arr(ff.fene_atoms, np.int32, C.c_int32), arr(ff.fene_param, np.float64, C.c_double)
    v = None
    if velocities and system.v is not None:
        v = np.ascontiguousarray(system.v, np.float64).reshape(-1)
    _check(lib().capsgpu_write_lammps(os.fsencode(path), C.byref(S), _p(v), title.encode()))


def write_checkpoint(path, system, step, total):
    """a checkpoint as caps md writes one (cli/main.cpp:7820-7846): the state with velocities as a LAMMPS data file whose
    title says the step, and the run settings beside it, each written whole (a temporary file, then renamed)"""
    import json
    write_lammps(path, system, "checkpoint step %d of %d" % (step, total))
    ff = system.ff
    sp = _settings_path(path)
    with open(sp + ".tmp", "w") as f:
        json.dump({"step": int(step), "total": int(total), "lj_shift": bool(ff.lj_shift), "cutoff": float(ff.cutoff),
                   "coul14": float(ff.coul14), "lj96": bool(ff.lj96), "exclude13": _excludes_13(ff)}, f)
    os.replace(sp + ".tmp", sp)


def checkpoint_step(system):
    """the step a checkpoint was written at, from its title (cli/main.cpp:7810-7816); raises ValueError for another file
    or one without velocities"""
    at = system.title.find("checkpoint step ")
    if at < 0:
        raise ValueError("not a checkpoint: its title does not say the step")
    if system.v is None or len(system.v) != system.ff.n:
        raise ValueError("the checkpoint carries no velocities")
    return int(system.title[at + 16:].split()[0])


def resume_options(system, options, total_steps):
    """the options that continue a run from a checkpoint read with read_lammps: total_steps is the run's whole length,
    the checkpoint's step becomes the step offset and its velocities are kept (cli/main.cpp:7809-7820)"""
    import copy
    done = checkpoint_step(system)
    if done >= total_steps:
        raise ValueError("the checkpoint is at step %d: the total number of steps must be larger" % done)
    o = copy.copy(options)
    o.step_offset, o.steps, o.new_velocities = done, total_steps - done, False
    return o


_MINIMISERS = {"sd": 0, "steepest": 0, "cg": 1, "lbfgs": 2, "l-bfgs": 2, "fire": 3}
_STOPS = {0: "iterations", 1: "force", 2: "energy", 3: "line search"}


@dataclass
class RelaxOptions:
    """caps::RelaxOptions (core/include/caps/relax.hpp:28-41), the minimiser's fields"""
    method: str = "lbfgs"             # sd, cg, lbfgs, fire
    ftol: float = 0.5                 # kcal/mol/A
    etol: float = 1e-8
    max_iterations: int = 5000
    fixed: np.ndarray = None          # per atom: held in place


@dataclass
class RelaxStage:
    """caps::RelaxStage (core/include/caps/relax.hpp:86-91)"""
    iterations: int
    evaluations: int
    energy: float
    fmax: float
    stopped_by: str


def minimise(ev, system, o=None):
    """caps::minimise (core/src/relax.cpp:65-282) on the GPU with the evaluator's terms: system.x is updated"""
    o = o or RelaxOptions()
    if o.method.lower() not in _MINIMISERS:
        raise ValueError("unknown minimiser '%s' (sd, cg, lbfgs, fire)" % o.method)
    r = ROpts()
    r.method, r.ftol, r.etol, r.max_iterations = _MINIMISERS[o.method.lower()], o.ftol, o.etol, o.max_iterations
    x = np.ascontiguousarray(system.x, np.float64).reshape(-1).copy()
    cell = np.ascontiguousarray(system.cell, np.float64)
    fixed = None if o.fixed is None else np.ascontiguousarray(o.fixed, np.uint8)
    st = RStage()
    _check(lib().capsgpu_minimise(ev._h, _p(x), _p(cell), C.byref(r), _p(fixed), C.byref(st)))
    system.x = x.reshape(-1, 3)
    return RelaxStage(st.iterations, st.evaluations, st.energy, st.fmax, _STOPS.get(st.stopped_by, "?"))


# ---- the glass-transition scan (caps::run_cooling, core/src/mechanics.cpp:1381-1487) --------------------------------
@dataclass
class CoolingOptions:
    """caps::CoolingOptions (core/include/caps/mechanics.hpp:323-343), the fields this path reads"""
    t_start: float = 500.0
    t_end: float = 200.0
    t_step: float = 20.0
    ps_per_step: float = 100.0
    equilibrate_ps: float = -1.0      # < 0: one hold
    average_from: float = 0.5
    dt: float = 1.0
    tau_t: float = 100.0
    tau_p: float = 1000.0
    pressure: float = 1.0
    compressibility: float = 4.5e-5
    barostat: str = "c-rescale"
    barostat_every: int = 10
    seed: int = 1
    new_velocities: bool = False
    property: str = "volume"          # volume (specific volume) or potential (potential energy per atom)
    fit: str = "bilinear"             # bilinear (hinge free) or ranges (glassy and rubbery lines)
    glassy_max: float = 0.0
    rubbery_min: float = 0.0
    energy: EnergyOptions = field(default_factory=EnergyOptions)
    progress: object = None           # callable(row, index, total) -> bool


TG_NOTES = {0: "", 1: "too few temperatures for the fit", 2: "the two slopes differ by less than 10 %: no clear transition in this range",
            3: "the hinge sits at the edge of the range: extend the temperatures", 4: "the glassy range must end below where the rubbery range starts",
            5: "each range needs at least two temperatures", 6: "the two lines are parallel: no crossing",
            7: "a range with two temperatures fixes its line exactly: no error estimate", 8: "the lines cross outside the gap between the ranges: no Tg"}


@dataclass
class TgFit:
    """caps::BilinearFit (core/include/caps/mechanics.hpp:305-313)"""
    ok: bool
    tg: float
    tg_err: float
    slope_low: float
    slope_high: float
    value_at_tg: float
    alpha_low: float
    alpha_high: float
    rss: float
    note: str


@dataclass
class CoolingResult:
    points: list                      # dicts: temperature, density, density_err, specific_volume, potential
    fit: TgFit


def _tgfit(f):
    return TgFit(bool(f.ok), f.tg, f.tg_err, f.slope_low, f.slope_high, f.value_at_tg, f.alpha_low, f.alpha_high, f.rss, TG_NOTES.get(f.note, ""))


def _points(arr, n):
    return [dict(temperature=p.temperature, density=p.density, density_err=p.density_err, specific_volume=p.specific_volume,
                 potential=p.potential) for p in arr[:n]]


def _copts(o):
    c = COpts()
    c.t_start, c.t_end, c.t_step, c.ps_per_step = o.t_start, o.t_end, o.t_step, o.ps_per_step
    c.equilibrate_ps, c.average_from, c.dt, c.tau_t = o.equilibrate_ps, o.average_from, o.dt, o.tau_t
    c.tau_p, c.pressure, c.compressibility = o.tau_p, o.pressure, o.compressibility
    if o.barostat.lower() not in _BAROSTATS:
        raise ValueError("unknown barostat '%s' (none, berendsen, c-rescale)" % o.barostat)
    c.barostat, c.barostat_every, c.seed = _BAROSTATS[o.barostat.lower()], o.barostat_every, o.seed
    c.new_velocities = int(o.new_velocities)
    c.property = {"volume": 0, "potential": 1}[o.property]
    c.fit = {"bilinear": 0, "ranges": 1}[o.fit]
    c.glassy_max, c.rubbery_min = o.glassy_max, o.rubbery_min
    return c


def _ladder_len(o):
    import math
    return int(math.floor(abs(o.t_end - o.t_start) / o.t_step + 1e-9)) + 1


def run_cooling(system, o, evaluator=None):
    """caps::run_cooling (core/src/mechanics.cpp:1381-1487) on the GPU: stepwise cooling of one system (its state left
    at the end of the scan) and the Tg fit"""
    ev = evaluator or Evaluator(system.ff, o.energy)
    ev.set_state(system.x, None if o.new_velocities or system.v is None else system.v, system.cell)
    c = _copts(o)
    if system.v is None:
        c.new_velocities = 1
    cap = _ladder_len(o) + 2
    pts = (CPoint * cap)()
    n = C.c_int32()
    f = CFit()
    cb = None
    if o.progress is not None:
        def _cb(row, index, total, user):
            return 1 if o.progress(_row(row.contents), index, total) else 0
        cb = COOLING_FN(_cb)
    _check(lib().capsgpu_run_cooling(ev._h, C.byref(c), C.cast(cb, C.c_void_p) if cb else None, None, C.cast(pts, C.c_void_p), cap,
                                     C.byref(n), C.byref(f)))
    x, v, cell = ev.get_state()
    system.x, system.v, system.cell = x, v, cell
    return CoolingResult(_points(pts, n.value), _tgfit(f))


def fit_tg(T, y, fit="bilinear", glassy_max=0.0, rubbery_min=0.0):
    """caps::fit_bilinear / fit_two_ranges (core/src/mechanics.cpp:1218-1379) of given points"""
    T = np.ascontiguousarray(T, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    f = CFit()
    _check(lib().capsgpu_tg_fit_points(_p(T), _p(y), len(T), {"bilinear": 0, "ranges": 1}[fit], glassy_max, rubbery_min, C.byref(f)))
    return _tgfit(f)


def batch_cooling(batch, systems, o, samples=5):
    """many independent replicas scanned together through the replica batch (Langevin at constant volume; Tg from the
    potential energy per atom): one CoolingResult per replica; the systems' states are updated"""
    nrep = len(batch.evaluators)
    if len(systems) != nrep:
        raise ValueError("one system per replica")
    for ev, s in zip(batch.evaluators, systems):
        ev.set_state(s.x, None if o.new_velocities or s.v is None else s.v, s.cell)
    c = _copts(o)
    if any(s.v is None for s in systems):
        c.new_velocities = 1
    cap = _ladder_len(o) + 2
    pts = (CPoint * (cap * nrep))()
    fits = (CFit * nrep)()
    n = C.c_int32()
    _check(lib().capsgpu_batch_cooling(batch._h, C.byref(c), samples, C.cast(pts, C.c_void_p), cap, C.byref(n), C.cast(fits, C.c_void_p)))
    out = []
    for r, (ev, s) in enumerate(zip(batch.evaluators, systems)):
        x, v, cell = ev.get_state()
        s.x, s.v, s.cell = x, v, cell
        out.append(CoolingResult(_points(pts[r * cap:(r + 1) * cap], n.value), _tgfit(fits[r])))
    return out


# ---- equilibration protocols (caps::equilibrate, core/src/equilibrate.cpp) -----------------------------------------
_ENSEMBLES = {"nvt": 0, "npt": 1, "nve": 2}


@dataclass
class Stage:
    """caps::Stage (core/include/caps/equilibrate.hpp:15-22), without the push-off force cap"""
    ensemble: str = "nvt"
    ps: float = 10.0
    t_start: float = 300.0
    t_end: float = -1.0
    pressure: float = 1.0
    label: str = ""


@dataclass
class ProtocolParams:
    """caps::ProtocolParams (core/include/caps/equilibrate.hpp:24-37)"""
    t_final: float = 300.0
    t_max: float = 600.0
    p_final: float = 1.0
    p_max: float = 49346.2
    time_scale: float = 1.0
    cycles: int = 3
    t_low: float = 300.0
    t_high: float = 600.0
    ramp_ps: float = 50.0
    hold_ps: float = 50.0


def protocol(name, params=None):
    """the stages of "larsen21" or "annealing" (core/src/equilibrate.cpp:43-75)"""
    p = params or ProtocolParams()
    pp = ProtoParams(p.t_final, p.t_max, p.p_final, p.p_max, p.time_scale, p.cycles, p.t_low, p.t_high, p.ramp_ps, p.hold_ps)
    n = C.c_int32()
    _check(lib().capsgpu_protocol(name.encode(), C.byref(pp), None, 0, C.byref(n)))
    arr = (EqStage * max(1, n.value))()
    _check(lib().capsgpu_protocol(name.encode(), C.byref(pp), C.cast(arr, C.c_void_p), n.value, C.byref(n)))
    names = {0: "nvt", 1: "npt", 2: "nve"}
    return [Stage(names[a.ensemble], a.ps, a.t_start, a.t_end, a.pressure, "%d" % (k + 1)) for k, a in enumerate(arr[:n.value])]


@dataclass
class EquilibrateOptions:
    """caps::EquilibrateOptions (core/include/caps/equilibrate.hpp:65-89), the staged part"""
    stages: list = field(default_factory=list)
    dt: float = 1.0
    tau_t: float = 100.0
    tau_p: float = 1000.0
    compressibility: float = 4.5e-5
    thermostat: str = "bussi"
    barostat: str = "c-rescale"
    barostat_every: int = 10
    seed: int = 1
    thermo_ps: float = 0.0
    energy: EnergyOptions = field(default_factory=EnergyOptions)
    progress: object = None           # callable(stage_index, stages, row) -> bool


EQ_FN = C.CFUNCTYPE(C.c_int32, C.POINTER(Thermo), C.c_int32, C.c_int32, C.c_void_p)


def equilibrate(system, o, evaluator=None):
    """caps::equilibrate (core/src/equilibrate.cpp:188-262) on the GPU: the stages one after another; returns one summary
    dict per stage (means over its second half) and leaves the system at the end"""
    ev = evaluator or Evaluator(system.ff, o.energy)
    ev.set_state(system.x, system.v, system.cell)
    st = (EqStage * len(o.stages))()
    for k, s in enumerate(o.stages):
        st[k] = EqStage(_ENSEMBLES[s.ensemble.lower()], s.ps, s.t_start, s.t_end, s.pressure)
    eo = EqOpts(o.dt, o.tau_t, o.tau_p, o.compressibility, _THERMOSTATS[o.thermostat.lower()], _BAROSTATS[o.barostat.lower()],
                o.barostat_every, o.seed, o.thermo_ps)
    out = (EqSummary * len(o.stages))()
    cb = None
    if o.progress is not None:
        def _cb(row, index, total, user):
            return 1 if o.progress(index, total, _row(row.contents)) else 0
        cb = EQ_FN(_cb)
    _check(lib().capsgpu_equilibrate(ev._h, C.cast(st, C.c_void_p), len(o.stages), C.byref(eo), C.cast(cb, C.c_void_p) if cb else None,
                                     None, C.cast(out, C.c_void_p)))
    x, v, cell = ev.get_state()
    system.x, system.v, system.cell = x, v, cell
    return [dict(label=s.label, ps=r.ps, temperature=r.temperature, pressure=r.pressure, density=r.density, potential=r.potential,
                 density_end=r.density_end) for s, r in zip(o.stages, out)]

