"""ctypes declarations of include/ne_gpu.h."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

NE_FP32, NE_FP64 = 0, 1
NE_FLAG_COUNT = 1
NE_FLAG_TIME = 2
NE_FLAG_STAGE_TIME = 4
NE_FLAG_SORT = 8
NE_FLAG_FIXED_FILM = 16
NE_FLAG_SPATIAL = 32
NE_FLAG_WIDE_OFFSETS = 64
NE_FLAG_TEST_PLAN_STALL = 128
STAGES = ("megakernel", "plan", "regen", "trace", "medium", "shade", "shadow", "other", "guide", "train")
NE_MODE_AUTO, NE_MODE_MEGAKERNEL, NE_MODE_WAVEFRONT = 0, 1, 2
NE_BATCH_PERSISTENT, NE_BATCH_WAVEFRONT = 0, 1  # batch executors
NE_BATCH_LANES = 4
NE_MAT_GLOSSY, NE_MAT_EMITTER = 0, 1
NE_ENV_NONE, NE_ENV_CONSTANT, NE_ENV_SKY, NE_ENV_MAP = 0, 1, 2, 3
NE_TEST_AABB, NE_TEST_TRIANGLE, NE_TEST_SPHERE, NE_TEST_RECTANGLE = 0, 1, 2, 3
NE_TEST_HG_EVAL, NE_TEST_HG_SAMPLE, NE_TEST_GGX_D, NE_TEST_GGX_PDF, NE_TEST_GLOSSY, NE_TEST_HG_FWD = range(6)

_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_D = C.POINTER(C.c_double)


class SceneDesc(C.Structure):
    _fields_ = [
        ("tri_v", _F), ("tri_mat", _I), ("n_tri", C.c_uint32),
        ("sph", _F), ("sph_mat", _I), ("n_sph", C.c_uint32),
        ("rect", _F), ("rect_mat", _I), ("n_rect", C.c_uint32),
        ("mat", _F), ("n_mat", C.c_uint32),
        ("has_volume", C.c_int32),
        ("vol_xf", C.c_float * 16), ("vol_inv", C.c_float * 16),
        ("vol_res", C.c_int32 * 3),
        ("grid", _F),
        ("vol_mult", C.c_float), ("vol_sigma_s", C.c_float * 3), ("vol_sigma_a", C.c_float * 3), ("vol_g", C.c_float),
        ("maj_res", C.c_int32),
        ("has_sun", C.c_int32), ("sun_dir", C.c_float * 3), ("sun_E", C.c_float * 3),
        ("env_type", C.c_int32), ("env_c0", C.c_float * 3), ("env_c1", C.c_float * 3), ("env_c2", C.c_float * 3),
        ("vol_homogeneous", C.c_int32),
        ("env_map", _F), ("env_w", C.c_int32), ("env_h", C.c_int32),
        ("point", _F), ("point_mat", _I), ("n_point", C.c_uint32),
        ("maj_no_skip", C.c_int32),
    ]


class CameraDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("look_at", C.c_float * 3), ("up", C.c_float * 3),
                ("vfov", C.c_float), ("aspect", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32),
        ("x0", C.c_int32), ("y0", C.c_int32), ("x1", C.c_int32), ("y1", C.c_int32),
        ("spp", C.c_int32), ("max_bounces", C.c_int32),
        ("seed", C.c_uint64), ("sample_offset", C.c_uint32), ("job", C.c_uint32),
        ("precision", C.c_int32), ("flags", C.c_int32),
        ("samples_per_thread", C.c_int32), ("block_size", C.c_int32),
        ("occupancy", C.c_int32), ("mode", C.c_int32), ("pool_size", C.c_int32),
    ]


class Counters(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in
                ("paths", "ext_rays", "shadow_rays", "density_lookups", "majorant_cells", "bvh_nodes",
                 "prim_tests", "shadow_density_lookups", "shadow_majorant_cells", "shadow_bvh_nodes",
                 "shadow_prim_tests", "guide_queries")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class SceneInfo(C.Structure):
    _fields_ = [("bvh_nodes", C.c_uint32), ("bvh_depth", C.c_uint32), ("bvh_leaf_prims_max", C.c_uint32),
                ("maj_res", C.c_uint32), ("device_bytes", C.c_uint64), ("grid_max_density", C.c_float),
                ("pad", C.c_float)]


class NarvalError(RuntimeError):
    pass


# (name, restype, argtypes) for every entry point declared in include/ne_gpu.h
class Job(C.Structure):
    """ne_job: one render job of a batch."""
    _fields_ = [("scene", C.c_void_p), ("camera", CameraDesc), ("width", C.c_int32), ("height", C.c_int32),
                ("spp", C.c_int32), ("max_bounces", C.c_int32), ("seed", C.c_uint64), ("job", C.c_uint32),
                ("reserved", C.c_uint32), ("guide", C.c_void_p)]


class BatchConfig(C.Structure):
    """ne_batch_config: executor and, for the persistent kernel, the work-item shape."""
    _fields_ = [("tile_w", C.c_int32), ("tile_h", C.c_int32), ("samples_per_item", C.c_int32),
                ("executor", C.c_int32)]


class GuideConfig(C.Structure):
    """ne_guide_config."""
    _fields_ = [("log2_table", C.c_int32), ("base_res", C.c_int32), ("max_res", C.c_int32), ("batch", C.c_int32),
                ("lr_table", C.c_float), ("lr_mlp", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("guided_vertices", C.c_int32), ("guide_prob", C.c_float), ("seed", C.c_uint64),
                ("cache_res", C.c_int32), ("cache_dirs", C.c_int32)]


class GuideStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("samples", C.c_uint64), ("last_loss", C.c_float), ("pad", C.c_float)]


class SceneSettings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("bounces", C.c_int32),
                ("hdr", C.c_int32), ("realtime", C.c_int32)]


class LoadOptions(C.Structure):
    _fields_ = [("flags", C.c_int32), ("maj_res", C.c_int32), ("resources_dir", C.c_char_p)]


NE_LOAD_LENIENT = 1
NE_IMAGE_PPM16, NE_IMAGE_PNG, NE_IMAGE_EXR, NE_IMAGE_PFM = 0, 1, 2, 3

SIGNATURES = [
    ("ne_ctx_create", C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("ne_ctx_destroy", None, [C.c_void_p]),
    ("ne_last_error", C.c_char_p, [C.c_void_p]),
    ("ne_synchronize", C.c_int32, [C.c_void_p]),
    ("ne_sizeof_scene_desc", C.c_int32, []),
    ("ne_scene_upload", C.c_int32, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    ("ne_scene_release", None, [C.c_void_p]),
    ("ne_scene_get_info", C.c_int32, [C.c_void_p, C.POINTER(SceneInfo)]),
    ("ne_film_create", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("ne_film_release", None, [C.c_void_p]),
    ("ne_film_clear", C.c_int32, [C.c_void_p]),
    ("ne_film_read", C.c_int32, [C.c_void_p, _D]),
    ("ne_film_tonemap", C.c_int32, [C.c_void_p, C.c_int32, _F]),
    ("ne_render", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(CameraDesc), C.POINTER(RenderParams), C.c_void_p]),
    ("ne_render_noalloc", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(CameraDesc), C.POINTER(RenderParams),
                                      C.c_void_p]),
    ("ne_render_tile", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(CameraDesc), C.POINTER(RenderParams),
                                   C.c_int32, C.c_void_p, _F]),
    ("ne_get_kernel_time", C.c_int32, [C.c_void_p, _F]),
    ("ne_get_counters", C.c_int32, [C.c_void_p, C.POINTER(Counters)]),
    ("ne_get_stage_times", C.c_int32, [C.c_void_p, _F, C.POINTER(C.c_uint32)]),
    ("ne_li", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, _F, C.c_int32, C.c_uint64, C.c_uint32, _F]),
    ("ne_test_primitive", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _F, _F, _I, _F, _F]),
    ("ne_test_closest", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, _F, C.c_float, C.c_float, _F, _I, _I]),
    ("ne_test_transmittance", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, _F, _F, C.c_uint64, C.c_int32, _D]),
    ("ne_test_delta_track", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, _F, C.c_uint64, C.c_int32, _I, _D]),
    ("ne_test_density", C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, _F, _F]),
    ("ne_test_majorants", C.c_int32, [C.c_void_p, C.c_void_p, _F, C.c_int64]),
    ("ne_test_bsdf", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, _F, _F, _F, _F, _F]),
    ("ne_test_rng", C.c_int32, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, _F]),
    ("ne_bvh_build", C.c_int32, [C.POINTER(SceneDesc), _F, _U, _U, _U, _U]),
    ("ne_gen_fbm_grid", C.c_int32, [C.c_int32, C.c_uint64, C.c_float, _F]),
    ("ne_gen_cumulus_grid", C.c_int32, [C.c_int32, C.c_uint64, _F]),
    ("ne_gen_icosphere", C.c_int32, [C.c_int32, C.c_uint64, C.c_float, _F, _U, _U, _U]),
    ("ne_batch_create", C.c_int32, [C.c_void_p, C.POINTER(Job), C.c_int32, C.POINTER(C.c_void_p)]),
    ("ne_batch_release", None, [C.c_void_p]),
    ("ne_batch_autotune", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(BatchConfig), _F]),
    ("ne_batch_run", C.c_int32, [C.c_void_p, C.POINTER(BatchConfig), _F]),
    ("ne_batch_job_times", C.c_int32, [C.c_void_p, _F]),
    ("ne_batch_read", C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    ("ne_guide_create", C.c_int32, [C.c_void_p, C.POINTER(GuideConfig), C.POINTER(C.c_void_p)]),
    ("ne_render_guided", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(CameraDesc), C.POINTER(RenderParams),
                                     C.c_void_p, C.c_void_p, C.c_int32]),
    ("ne_guide_release", None, [C.c_void_p]),
    ("ne_guide_get_cache", C.c_int32, [C.c_void_p, _F, C.POINTER(C.c_uint64)]),
    ("ne_guide_reset", C.c_int32, [C.c_void_p]),
    ("ne_guide_get_stats", C.c_int32, [C.c_void_p, C.POINTER(GuideStats)]),
    ("ne_guide_eval", C.c_int32, [C.c_void_p, C.c_int32, _F, C.c_int32, _F]),
    ("ne_guide_train_step", C.c_int32, [C.c_void_p, C.c_int32, _F, C.c_int32, _F]),
    ("ne_guide_sizes", C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("ne_guide_get_params", C.c_int32, [C.c_void_p, _F, _F]),
    ("ne_guide_set_params", C.c_int32, [C.c_void_p, _F, _F]),
    ("ne_guide_get_grads", C.c_int32, [C.c_void_p, _F, _F]),
    ("ne_scene_file_read", C.c_int32, [C.c_char_p, C.POINTER(LoadOptions), C.POINTER(C.c_void_p)]),
    ("ne_scene_file_parse", C.c_int32, [C.c_char_p, C.c_size_t, C.POINTER(LoadOptions), C.POINTER(C.c_void_p)]),
    ("ne_scene_file_get", C.c_int32, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(CameraDesc),
                                      C.POINTER(SceneSettings)]),
    ("ne_scene_file_warnings", C.c_char_p, [C.c_void_p]),
    ("ne_scene_file_release", None, [C.c_void_p]),
    ("ne_vol_read", C.c_int32, [C.c_char_p, C.POINTER(C.c_int32), _F, C.c_uint64]),
    ("ne_image_write", C.c_int32, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, _F]),
]

_LIB = None


def lib_path() -> Path:
    return Path(__file__).resolve().parent / "libnarval_gpu.so"


def lib():
    """Loads libnarval_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not p.exists():
            raise NarvalError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `python narvalengine_amd/build.py`")
        L = C.CDLL(str(p))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.ne_sizeof_scene_desc() != C.sizeof(SceneDesc):
            raise NarvalError("ne_scene_desc layout mismatch between libnarval_gpu.so and abi.py")
        _LIB = L
    return _LIB


def check(status: int, ctx=None) -> None:
    if status != 0:
        msg = lib().ne_last_error(ctx)
        raise NarvalError(f"ne status {status}: {msg.decode() if msg else ''}")
