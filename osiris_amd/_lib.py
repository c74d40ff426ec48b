"""ctypes binding of libosiris_amd.so (the C ABI of include/osiris_amd.h).

The library is built in-tree (``osiris_amd/lib/libosiris_amd.so``).  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libosiris_amd.so")

OSR_OK = 0
OSR_BORIS = 0
OSR_VAY = 1
OSR_PATH_FUSED = 0
OSR_PATH_PLAIN = 1
OSR_BND_PERIODIC = 0
OSR_BND_OPEN = 1
OSR_BND_LINDMAN = 1
OSR_BND_VPML = 2

P = C.c_void_p
I = C.c_int
I64 = C.c_int64
D = C.c_double
F = C.c_float

# name -> (restype, argtypes); every symbol of include/osiris_amd.h
SIGNATURES = {
    "osr_last_error": (C.c_char_p, []),
    "osr_version": (I, []),
    "osr_device_count": (I, []),
    "osr_advance_deposit": (I, [I, I, I, P, P, P, P, I, I, F, P, P, P, P, P, P, P, P, D, P]),
    "osr_advance_deposit_noalloc": (I, [I, I, I, P, P, P, P, I, I, F, P, P, P, P, P, P, P, P, D, P]),
    "hip_vadvance_deposit_2d_s1_": (None, [P] * 17),
    "hip_vadvance_deposit_2d_s2_": (None, [P] * 17),
    "hip_vadvance_deposit_2d_s3_": (None, [P] * 17),
    "hip_vadvance_deposit_3d_s1_": (None, [P] * 17),
    "hip_vadvance_deposit_3d_s2_": (None, [P] * 17),
    "hip_vadvance_deposit_3d_s3_": (None, [P] * 17),
    "osr_yee_advance": (I, [I, P, P, P, P, D, P, P, P]),
    "osr_sim_create": (I, [P, P]),
    "osr_sim_destroy": (I, [P]),
    "osr_sim_add_species": (I, [P, F, I, P]),
    "osr_species_set": (I, [P, I, I64, P, P, P, P]),
    "osr_species_count": (I, [P, I, P]),
    "osr_species_get": (I, [P, I, P, P, P, P]),
    "osr_species_generate": (I, [P, I, P, D, P, P, C.c_uint64, I]),
    "osr_fields_set": (I, [P, P, P, P, P]),
    "osr_fields_get": (I, [P, P, P, P, P]),
    "osr_fields_fill": (I, [P, P, P]),
    "osr_sim_run": (I, [P, I]),
    "osr_sim_steps": (I, [P, P]),
    "osr_sim_history": (I, [P, I64, I64, P, P]),
    "osr_sim_gauss": (I, [P, P, P, P]),
    "osr_sim_relayouts": (I, [P, P]),
    "osr_sim_time_kernels": (I, [P, I, P]),
    "osr_sim_footprint": (I, [P, P]),
    "osr_sim_launch_config": (I, [P, P, P, P]),
    "osr_ens_create": (I, [P, P]),
    "osr_ens_destroy": (I, [P]),
    "osr_ens_submit": (I, [P, P, P]),
    "osr_ens_progress": (I, [P, P, P]),
    "osr_ens_wait": (I, [P]),
    "osr_ens_job_info": (I, [P, I64, P]),
    "osr_ens_job_history": (I, [P, I64, I64, I64, P, P]),
    "osr_ens_job_release": (I, [P, I64]),
    "osr_ens_usage": (I, [P, P]),
    "osr_smooth_setup": (I, [I, P, P, D, P, P]),
    "osr_smooth_vdf": (I, [I, P, I, P, P, P]),
    "osr_sim_set_smooth": (I, [P, P, P]),
    "osr_hist_append": (I, [P, I, P, I64, P, P, P, P, P]),
    "osr_sim_write_hist": (I, [P, P, P, I64, I64]),
    "osr_zdf_save_grid": (I, [P, I, P, P, P]),
    "osr_zdf_write_grid_file": (I, [P, P, I, P, P]),
    "osr_zdf_save_particles": (I, [P, P, P, P]),
    "osr_sim_save_field": (I, [P, I, P, P]),
    "osr_zpulse_wall_defaults": (None, [P]),
    "osr_sim_init_fields": (I, [P, I, C.c_float]),
    "osr_sim_add_zpulse_wall": (I, [P, P]),
    "osr_sim_zpulse_column": (I, [P, I, I, C.c_double, P, P, P]),
    "osr_zpulse_defaults": (None, [P]),
    "osr_sim_add_zpulse": (I, [P, P]),
    "osr_copy_probe": (I, [I64, I, P]),
}


class ZpulseWall(C.Structure):
    """osr_zpulse_wall"""
    _fields_ = [("a0", C.c_double), ("omega0", C.c_double), ("phase", C.c_double),
                ("pol", C.c_double), ("pol_type", C.c_int), ("propagation", C.c_int),
                ("tenv_type", C.c_int), ("per_type", C.c_int), ("tenv_fwhm", C.c_double),
                ("tenv_rise", C.c_double), ("tenv_flat", C.c_double), ("tenv_fall", C.c_double),
                ("tenv_duration", C.c_double), ("tenv_range", C.c_double),
                ("per_center", C.c_double * 2), ("per_w0", C.c_double), ("per_fwhm", C.c_double),
                ("per_focus", C.c_double), ("per_tem_mode", C.c_int * 2),
                ("launch_time", C.c_double), ("xmin", C.c_double * 3)]


class ZpulseBox(C.Structure):
    """osr_zpulse (in-box pulse)"""
    _fields_ = [("a0", C.c_double), ("omega0", C.c_double), ("phase", C.c_double),
                ("pol", C.c_double), ("pol_type", C.c_int), ("propagation", C.c_int),
                ("lon_type", C.c_int), ("per_type", C.c_int), ("lon_start", C.c_double),
                ("lon_fwhm", C.c_double), ("lon_rise", C.c_double), ("lon_flat", C.c_double),
                ("lon_fall", C.c_double), ("lon_duration", C.c_double), ("lon_x0", C.c_double),
                ("lon_range", C.c_double), ("per_center", C.c_double * 2), ("per_w0", C.c_double),
                ("per_fwhm", C.c_double), ("per_focus", C.c_double), ("per_tem_mode", C.c_int * 2),
                ("no_div_corr", C.c_int), ("xmin", C.c_double * 3)]


OSR_TENV = {"polynomial": 0, "sin2": 1, "gaussian": 2}
OSR_PER = {"plane": 0, "gaussian": 1, "hermite": 1}
OSR_PROP = {"forward": 0, "backward": 1}


class SimDesc(C.Structure):
    _fields_ = [("dims", C.c_int), ("nx", C.c_int * 3), ("dx", C.c_double * 3),
                ("dt", C.c_double), ("order", C.c_int), ("tile", C.c_int * 3),
                ("path", C.c_int), ("kernel", C.c_int), ("xbnd", C.c_int),
                ("reorder", C.c_int), ("vpml_size", C.c_int),
                ("vpml_diffuse", C.c_int * 2), ("push_waves", C.c_int)]


class SmoothParams(C.Structure):
    """osr_smooth_params: the nl_smooth namelist, one entry per direction."""
    _fields_ = [("type", C.c_int * 3), ("order", C.c_int * 3),
                ("swfj", ((C.c_double * 3) * 15) * 3), ("laserkdx_omega0", C.c_double),
                ("laserkdx_dir", C.c_int), ("reserved", C.c_int),
                ("digital_A", C.c_double * 3), ("digital_fc", C.c_double * 3)]


OSR_SMOOTH = {"none": 0, "custom": 1, "binomial": 2, "compensated": 3, "5pass": 4, "tri6": 5,
              "tri8": 6, "laserkdx": 10, "digital": 11}


class EnsDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_active", C.c_int), ("chunk_steps", C.c_int),
                ("reserved", C.c_int * 5)]


class JobSpecies(C.Structure):
    _fields_ = [("rqm", C.c_float), ("pusher", C.c_int), ("ppc", C.c_int * 3),
                ("random_pos", C.c_int), ("density", C.c_double), ("ufl", C.c_double * 3),
                ("uth", C.c_double * 3), ("seed", C.c_uint64)]


class JobDesc(C.Structure):
    _fields_ = [("sim", SimDesc), ("nsteps", C.c_int64), ("e0", C.c_double * 3),
                ("b0", C.c_double * 3), ("nspecies", C.c_int), ("reserved", C.c_int),
                ("species", JobSpecies * 8)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("nspecies", C.c_int), ("steps", C.c_int64),
                ("particles", C.c_int64), ("footprint", C.c_int64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


OSR_JOB_PENDING, OSR_JOB_ACTIVE, OSR_JOB_DONE, OSR_JOB_FAILED = 0, 1, 2, 3


class OsrError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OsrError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int):
    if status != OSR_OK:
        msg = lib().osr_last_error().decode(errors="replace")
        raise OsrError(f"osiris_amd error {status}: {msg}")
