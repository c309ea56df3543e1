"""ctypes bindings of libretrocam_hip.so (C ABI in include/retrocam_hip.h) and
librc_host.so (include/rc_host.h).

The HIP library is loaded from retrocam_amd/lib/ only; a missing library raises
at import of the binding (there is no CPU fallback for the render path).
"""
from __future__ import annotations

import ctypes as C
import os

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HIP_LIB_PATH = os.path.join(LIBDIR, "libretrocam_hip.so")
HOST_LIB_PATH = os.path.join(LIBDIR, "librc_host.so")

# status codes
RC_OK = 0
RC_ERR_INVALID_ARG = 1
RC_ERR_NO_DEVICE = 2
RC_ERR_HIP = 3
RC_ERR_OOM = 4
RC_ERR_STATE = 5
RC_ERR_LIMIT = 6

# buffers
RC_BUF_ACCUM = 0
RC_BUF_GBUFFER = 1
RC_BUF_GALBEDO = 2
RC_BUF_VARIANCE = 3
RC_BUF_PRIMARY_ID = 4
RC_BUF_GBUFFER_ID = 5
RC_BUF_ACCUM64 = 6
RC_BUF_POST_HDR = 7
RC_BUF_DISPLAY = 8

RC_ACCUM_FP32 = 0
RC_ACCUM_FP64 = 1

RC_PT_MEGAKERNEL = 0
RC_PT_WAVEFRONT = 1
RC_PT_LANES = 2
RC_LANE_VARIANTS = 11

RC_DENOISE_ATROUS = 0
RC_POST_FUSED = 0
RC_POST_PER_PASS = 1
RC_NET_KERNELS_TILED = 0
RC_NET_KERNELS_DIRECT = 1
RC_NET_KERNELS_SINGLE_GROUP = 2
RC_DENOISE_NET = 1

F3 = C.c_float * 3


class RcCamera(C.Structure):
    _fields_ = [
        ("camPos", F3), ("camRight", F3), ("camUp", F3), ("camForward", F3),
        ("focalDist", C.c_float), ("lensRadius", C.c_float), ("fovTanHalf", C.c_float),
        ("apertureBlades", C.c_int32), ("saStrength", C.c_float), ("lcaDelta", F3),
        ("width", C.c_int32), ("height", C.c_int32),
    ]


class RcPtParams(C.Structure):
    _fields_ = [("neeEnabled", C.c_int32), ("neeFireflyClamp", C.c_float), ("maxBounces", C.c_int32)]


_F = C.c_float
_I = C.c_int32


class RcPostParams(C.Structure):
    _fields_ = [
        ("exposure", _F), ("atrousEnabled", _I), ("atrousIterations", _I), ("atrousMaxSpp", _I),
        ("atrousSigmaColor", _F), ("atrousSigmaNormal", _F), ("atrousSigmaDepth", _F),
        ("enable", _I * 17),
        ("lumaBlurSigma", _F), ("chromaResSigma", _F),
        ("colorTempBias", _F), ("colorHueRot", _F), ("colorSaturation", _F), ("colorDriftSpeed", _F),
        ("drBlackLift", _F), ("drShadowCrush", _F), ("drHighlightKnee", _F), ("drHighlightClip", _F),
        ("ccdNoiseLumaBase", _F), ("ccdNoiseChromaScale", _F),
        ("ccdSmearThreshold", _F), ("ccdSmearIntensity", _F), ("ccdSmearLength", _I),
        ("chromaBleedFactor", _F),
        ("timebaseAmplitudePx", _F), ("timebaseFreq", _F), ("timebaseSpeed", _F),
        ("dotCrawlIntensity", _F), ("dotCrawlSubcarrier", _F), ("dotCrawlEdgeThresh", _F),
        ("tapeAge", _F), ("dropoutBright", _F), ("dropoutBurstRate", _F),
        ("headSwitchLines", _I), ("headSwitchJitter", _F), ("headSwitchRipple", _F), ("headSwitchChroma", _F),
        ("trackingSeverity", _F), ("trackingMaxDisplacePx", _F), ("trackingFringeWidthPx", _F),
        ("edgeEnhanceAmount", _F), ("edgeEnhanceCoreThreshold", _F),
        ("interlaceFieldOffsetPx", _F), ("interlaceCombStrength", _F), ("interlaceCombEdgeThresh", _F),
        ("interlaceLineWeighting", _F),
    ]


class RcView(C.Structure):
    _fields_ = [("scene", C.c_int32), ("nSamples", C.c_int32), ("cam", RcCamera), ("params", RcPtParams)]


RC_VIEWS_RESET = 1
RC_VIEWS_FP64 = 2
RC_MAX_SCENE_SLOTS = 64


class RcTuning(C.Structure):
    _fields_ = [("laneThreshold", C.c_int32), ("laneVariant", C.c_int32), ("laneBufferMiB", C.c_int32),
                ("laneTileWidth", C.c_int32), ("ptMode", C.c_int32)]


class RcTemporal(C.Structure):
    _fields_ = [("time", _F), ("agcGain", _F), ("frameIndex", _I)]


class RcsInfo(C.Structure):
    _fields_ = [
        ("nTris", C.c_int32), ("nNodes", C.c_int32), ("nMats", C.c_int32), ("nLights", C.c_int32),
        ("totalLightPower", C.c_float), ("maxDepth", C.c_int32), ("maxLeafSize", C.c_int32),
        ("nLeaves", C.c_int32),
    ]


class RctSettings(C.Structure):
    _fields_ = [("focusDistM", _F), ("afSpeed", _F), ("agcMaxGain", _F), ("agcSpeed", _F)]


class RctState(C.Structure):
    _fields_ = [("agcGain", _F), ("focalDist", _F), ("time", _F), ("agcGainTarget", _F),
                ("focalDistTarget", _F), ("whiteBalance", _F * 9)]


RCK_STEP, RCK_LINEAR, RCK_BEZIER = 0, 1, 2
RCS_OBJ_BOX, RCS_OBJ_SPHERE = 0, 1
RCS_JSON_OK, RCS_JSON_ERR_ARG, RCS_JSON_ERR_OOM, RCS_JSON_ERR_STRING = 0, 1, 2, 3


class RcsJsonObject(C.Structure):
    _fields_ = [("type", C.c_int32), ("pos", _F * 3), ("scale", _F * 3), ("rot", _F * 3),
                ("stacks", C.c_int32), ("slices", C.c_int32), ("materialIndex", C.c_int32)]

_FP = C.POINTER(C.c_float)
_VP = C.c_void_p

# name -> (restype, argtypes); the full exported surface of include/retrocam_hip.h
HIP_SYMBOLS = {
    "rc_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rc_destroy": (C.c_int, [_VP]),
    "rc_last_error": (C.c_char_p, [_VP]),
    "rc_stream": (C.c_void_p, [_VP]),
    "rc_sync": (C.c_int, [_VP]),
    "rc_upload_scene": (C.c_int, [_VP, _FP, C.c_int32, _FP, C.c_int32, _FP, C.c_int32, _FP, C.c_int32,
                                  C.c_float]),
    "rc_render_pass": (C.c_int, [_VP, C.POINTER(RcCamera), C.POINTER(RcPtParams)]),
    "rc_render_passes": (C.c_int, [_VP, C.POINTER(RcCamera), C.POINTER(RcPtParams), C.c_int32]),
    "rc_reset": (C.c_int, [_VP]),
    "rc_total_samples": (C.c_int32, [_VP]),
    "rc_set_accum_mode": (C.c_int, [_VP, C.c_int32]),
    "rc_set_pt_mode": (C.c_int, [_VP, C.c_int32]),
    "rc_read": (C.c_int, [_VP, C.c_int32, _VP, C.c_size_t]),
    "rc_buffer_bytes": (C.c_size_t, [_VP, C.c_int32]),
    "rc_sppm_reset": (C.c_int, [_VP, C.c_float]),
    "rc_sppm_set_precision": (C.c_int, [_VP, C.c_int32]),
    "rc_sppm_trace": (C.c_int, [_VP, C.c_int32, C.c_int32]),
    "rc_sppm_gather": (C.c_int, [_VP, C.POINTER(RcCamera)]),
    "rc_sppm_update_radius": (C.c_int, [_VP, C.c_float]),
    "rc_sppm_get_state": (C.c_int, [_VP, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "rc_sppm_read": (C.c_int, [_VP, _FP, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]),
    "rc_post_defaults": (None, [C.POINTER(RcPostParams)]),
    "rc_post": (C.c_int, [_VP, C.POINTER(RcPostParams), C.POINTER(RcTemporal)]),
    "rc_post_image": (C.c_int, [_VP, _FP, C.c_int32, C.c_int32, C.POINTER(RcPostParams), C.POINTER(RcTemporal)]),
    "rc_set_post_mode": (C.c_int, [_VP, C.c_int32]),
    "rc_post_launches": (C.c_int, [_VP, C.POINTER(C.c_int32)]),
    "rc_post_image_noalloc": (C.c_int, [_VP, _VP, C.c_int32, C.c_int32, C.POINTER(RcPostParams),
                                        C.POINTER(RcTemporal), _VP, _VP]),
    "rc_upload_scene_slot": (C.c_int, [_VP, C.c_int32, _FP, C.c_int32, _FP, C.c_int32, _FP, C.c_int32, _FP,
                                       C.c_int32, C.c_float]),
    "rc_views_render": (C.c_int, [_VP, C.POINTER(RcView), C.c_int32, C.c_int32]),
    "rc_views_post": (C.c_int, [_VP, C.c_int32, C.POINTER(RcPostParams), C.POINTER(RcTemporal)]),
    "rc_views_buffer_bytes": (C.c_size_t, [_VP, C.c_int32, C.c_int32]),
    "rc_views_read": (C.c_int, [_VP, C.c_int32, C.c_int32, _VP, C.c_size_t]),
    "rc_views_total_samples": (C.c_int32, [_VP, C.c_int32]),
    "rc_net_param_count": (C.c_int32, []),
    "rc_net_init": (C.c_int, [_VP, C.c_uint32]),
    "rc_net_set_params": (C.c_int, [_VP, _FP, C.c_int32]),
    "rc_net_get_params": (C.c_int, [_VP, _FP, C.c_int32]),
    "rc_net_set_dataset": (C.c_int, [_VP, _FP, _FP, C.c_int32, C.c_int32, C.c_int32]),
    "rc_net_train": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_float, C.c_uint32, _FP]),
    "rc_net_grad": (C.c_int, [_VP, _FP, _FP, C.c_int32, C.c_int32, C.c_int32, _FP, _FP]),
    "rc_net_forward": (C.c_int, [_VP, _FP, C.c_int32, C.c_int32, C.c_int32, _FP]),
    "rc_set_denoiser": (C.c_int, [_VP, C.c_int32]),
    "rc_export_frame": (C.c_int, [_VP, C.c_char_p, C.POINTER(C.c_uint8)]),
    "rc_display_size": (C.c_int, [_VP, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "rc_encode_png": (C.c_int, [C.POINTER(C.c_uint8), C.c_int32, C.c_int32, C.c_int32, C.c_char_p]),
    "rc_frame_capture_begin": (C.c_int, [_VP]),
    "rc_frame_capture_end": (C.c_int, [_VP]),
    "rc_frame_replay": (C.c_int, [_VP, C.POINTER(RcTemporal)]),
    "rc_frame_release": (C.c_int, [_VP]),
    "rc_frame_set_camera": (C.c_int, [_VP, C.POINTER(RcCamera)]),
    "rc_set_net_flips": (C.c_int, [_VP, C.c_int32]),
    "rc_set_net_kernels": (C.c_int, [_VP, C.c_int32]),
    "rc_get_tuning": (C.c_int, [_VP, C.POINTER(RcTuning)]),
    "rc_set_tuning": (C.c_int, [_VP, C.POINTER(RcTuning)]),
    "rc_autotune": (C.c_int, [_VP, C.POINTER(RcCamera), C.POINTER(RcPtParams), C.c_int32, C.POINTER(RcTuning),
                              C.POINTER(C.c_float)]),
    "rc_debug_counters": (C.c_int, [_VP, C.c_int32, C.POINTER(C.c_uint64), C.c_int32]),
    "rc_debug_pcg": (C.c_int, [_VP, C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.POINTER(C.c_uint32)]),
    "rc_debug_rays": (C.c_int, [_VP, _FP, C.c_int32, C.POINTER(C.c_int32), _FP, C.POINTER(C.c_int32)]),
    "rc_debug_recip_check": (C.c_int, [_VP, C.POINTER(C.c_uint64)]),
}

HOST_SYMBOLS = {
    "rcs_new": (C.c_void_p, []),
    "rcs_free": (None, [_VP]),
    "rcs_add_material": (C.c_int32, [_VP, _FP]),
    "rcs_add_box": (C.c_int32, [_VP] + [C.c_float] * 6 + [C.c_int32]),
    "rcs_add_sphere": (C.c_int32, [_VP] + [C.c_float] * 4 + [C.c_int32] * 3),
    "rcs_add_quad": (C.c_int32, [_VP, _FP, C.c_int32]),
    "rcs_rotate_since": (C.c_int32, [_VP, C.c_int32] + [C.c_float] * 6),
    "rcs_add_triangles": (C.c_int32, [_VP, _FP, C.c_int32]),
    "rcs_triangle_count": (C.c_int32, [_VP]),
    "rcs_build": (C.c_int32, [_VP]),
    "rcs_get_info": (C.c_int32, [_VP, C.POINTER(RcsInfo)]),
    "rcs_copy_tris": (C.c_int32, [_VP, _FP]),
    "rcs_copy_nodes": (C.c_int32, [_VP, _FP]),
    "rcs_copy_mats": (C.c_int32, [_VP, _FP]),
    "rcs_copy_lights": (C.c_int32, [_VP, _FP]),
    "rcs_orbit_basis": (None, [C.c_float] * 6 + [_FP] * 4),
    "rcs_lookat_basis": (None, [_FP] * 5),
    "rcs_tan_half_fov": (C.c_float, [C.c_float]),
    "rct_init": (None, [C.POINTER(RctState)]),
    "rct_update": (None, [C.POINTER(RctState), C.c_float, C.POINTER(RctSettings)]),
    "rct_reset_for_video": (None, [C.POINTER(RctState), C.POINTER(RctSettings)]),
    "rct_update_for_video_frame": (None, [C.POINTER(RctState), C.c_float, C.c_float, C.POINTER(RctSettings)]),
    "rct_run_video": (C.c_int32, [C.POINTER(RctState), C.POINTER(RctSettings), _FP, C.c_float, C.c_int32,
                                  _FP, _FP, _FP]),
    "rcj_frame_dt": (C.c_float, [C.c_float]),
    "rcj_total_frames": (C.c_int32, [C.c_float, C.c_float]),
    "rcj_canonical_time": (C.c_float, [C.c_int32, C.c_float]),
    "rcs_json_parse": (C.c_void_p, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int32)]),
    "rcs_json_free": (None, [_VP]),
    "rcs_json_material_count": (C.c_int32, [_VP]),
    "rcs_json_object_count": (C.c_int32, [_VP]),
    "rcs_json_get_material": (C.c_int32, [_VP, C.c_int32, _FP, C.c_char_p, C.c_int32]),
    "rcs_json_get_object": (C.c_int32, [_VP, C.c_int32, C.POINTER(RcsJsonObject), C.c_char_p, C.c_int32]),
    "rcs_add_object": (C.c_int32, [_VP, C.POINTER(RcsJsonObject)]),
    "rcs_json_build": (C.c_int32, [_VP, _VP]),
    "rck_new": (C.c_void_p, []),
    "rck_free": (None, [_VP]),
    "rck_add": (C.c_int32, [_VP, C.c_float, C.c_float, C.c_int32]),
    "rck_remove": (C.c_int32, [_VP, C.c_int32]),
    "rck_count": (C.c_int32, [_VP]),
    "rck_get": (C.c_int32, [_VP, C.c_int32, _FP, _FP, C.POINTER(C.c_int32)]),
    "rck_evaluate": (C.c_float, [_VP, C.c_float, C.c_float]),
    "rck_evaluate_many": (C.c_int32, [_VP, _FP, C.c_int32, C.c_float, _FP]),
}


def _bind(lib: C.CDLL, table: dict) -> C.CDLL:
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_hip = None
_host = None


def hip_lib() -> C.CDLL:
    """The loaded libretrocam_hip.so; raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError(
                f"{HIP_LIB_PATH} is missing: build it with `python -m retrocam_amd.build` "
                "(the render path has no CPU fallback)")
        _hip = _bind(C.CDLL(HIP_LIB_PATH), HIP_SYMBOLS)
    return _hip


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError(f"{HOST_LIB_PATH} is missing: build it with `python -m retrocam_amd.build`")
        _host = _bind(C.CDLL(HOST_LIB_PATH), HOST_SYMBOLS)
    return _host


def fptr(a) -> "C.POINTER(C.c_float)":
    return a.ctypes.data_as(_FP)
