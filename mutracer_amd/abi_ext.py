"""ctypes structures and signatures of the scene/render entry points of include/mtx.h."""
import ctypes as C


class MeshDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32),
                ("reserved", C.c_uint32), ("positions", C.c_void_p), ("normals", C.c_void_p),
                ("indices", C.c_void_p), ("uvs", C.c_void_p), ("uvs1", C.c_void_p)]


class TextureDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("storage", C.c_uint32), ("flip_y", C.c_uint32), ("uv_channel", C.c_uint32),
                ("wrap_s", C.c_uint32), ("wrap_t", C.c_uint32), ("uv_scale", C.c_float * 2),
                ("uv_offset", C.c_float * 2), ("pixels", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("mesh_count", C.c_uint32), ("meshes", C.c_void_p),
                ("instance_count", C.c_uint32), ("instances", C.c_void_p), ("material_count", C.c_uint32),
                ("materials", C.c_void_p), ("light_count", C.c_uint32), ("lights", C.c_void_p),
                ("env_width", C.c_uint32), ("env_height", C.c_uint32), ("env_rgba", C.c_void_p),
                ("blas_bins", C.c_uint32), ("tlas_bins", C.c_uint32), ("flags", C.c_uint32),
                ("texture_count", C.c_uint32), ("textures", C.c_void_p)]


class TargetSettings(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("bounces", C.c_uint32), ("direct_light_candidates", C.c_uint32),
                ("sampling_pattern", C.c_uint32), ("sequence_seed", C.c_uint32), ("max_samples", C.c_uint32),
                ("sample_sequence_offset", C.c_uint32), ("clamping", C.c_float), ("kernel", C.c_uint32),
                ("max_paths_in_flight", C.c_uint32)]


class CameraDesc(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("projection_inverse", C.c_float * 16),
                ("view_inverse", C.c_float * 16), ("aperture_radius", C.c_float), ("focus_distance", C.c_float)]


class WavefrontTuning(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_paths_in_flight", C.c_uint32), ("trace_kernel", C.c_uint32),
                ("lds_stack", C.c_uint32), ("trace_oversubscription", C.c_uint32),
                ("shade_oversubscription", C.c_uint32), ("shade_waves_per_simd", C.c_uint32),
                ("use_graphs", C.c_uint32), ("graph_replays", C.c_uint32), ("graph_captures", C.c_uint32),
                ("tail_paths", C.c_uint32), ("queue_sort", C.c_uint32), ("lds_stack_closest", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_paths_per_round", C.c_uint32), ("max_jobs_per_round", C.c_uint32),
                ("max_samples_per_job_round", C.c_uint32), ("hbm_budget_bytes", C.c_uint64)]


class JobInfo(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("state", C.c_uint32), ("requested_samples", C.c_uint32),
                ("submitted_samples", C.c_uint32), ("completed_samples", C.c_uint32), ("base_samples", C.c_uint32),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("complete_ms", C.c_double)]


class RenderSettings(C.Structure):
    """mtx_render_settings: the 64-byte mt_render_settings record."""
    _fields_ = [("struct_size", C.c_uint32), ("sampling_pattern", C.c_uint32), ("seed", C.c_uint32),
                ("max_samples", C.c_uint32), ("bounces", C.c_uint32), ("direct_light_candidates", C.c_uint32),
                ("bevel_samples", C.c_uint32), ("sample_sequence_offset", C.c_uint32), ("clamping", C.c_double),
                ("transmissive_output_alpha_min", C.c_double), ("caustic_flags", C.c_uint32),
                ("light_photons", C.c_uint32), ("photon_map_photons", C.c_uint32),
                ("photon_map_radius", C.c_float)]


class ContextSnapshot(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("samples", C.c_uint32), ("state", C.c_uint32), ("last_commit", C.c_uint32)]


class JobSnapshot(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("state", C.c_uint32), ("requested_samples", C.c_uint32),
                ("submitted_samples", C.c_uint32), ("completed_samples", C.c_uint32), ("base_samples", C.c_uint32)]


class AdaptiveOptions(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_samples", C.c_uint32), ("min_samples", C.c_uint32),
                ("tile_size", C.c_uint32), ("absolute_error", C.c_double), ("relative_error", C.c_double),
                ("confidence", C.c_double)]


def declare(L):
    u32, i32, u64, vp, sz = C.c_uint32, C.c_int32, C.c_uint64, C.c_void_p, C.c_size_t
    P = C.POINTER
    sig = {
        "mtx_scene_create": (i32, [vp, P(u64), C.c_char_p, sz]),
        "mtx_scene_set_transforms": (i32, [u64, vp, u32, vp, u32, C.c_char_p, sz]),
        "mtx_scene_info": (i32, [u64, vp]),
        "mtx_scene_update_mesh": (i32, [u64, u32, vp, vp, u32, P(u32), C.c_char_p, sz]),
        "mtx_scene_destroy": (i32, [u64]),
        "mtx_target_create": (i32, [u64, vp, vp, P(u64), C.c_char_p, sz]),
        "mtx_target_render": (i32, [u64, u32, vp]),
        "mtx_target_render_counted": (i32, [u64, u32, vp, vp]),
        "mtx_target_render_timed": (i32, [u64, u32, vp, vp]),
        "mtx_target_synchronize": (i32, [u64]),
        "mtx_target_read_pixels": (i32, [u64, u32, vp, sz, P(sz)]),
        "mtx_target_read_accumulators": (i32, [u64, vp, sz, vp, sz]),
        "mtx_target_stats": (i32, [u64, vp]),
        "mtx_target_set_camera": (i32, [u64, vp]),
        "mtx_target_reset": (i32, [u64]),
        "mtx_target_destroy": (i32, [u64]),
        "mtx_target_set_tuning": (i32, [u64, vp]),
        "mtx_target_plan_adaptive": (i32, [u64, vp, vp, vp, sz]),
        "mtx_target_set_sample_targets": (i32, [u64, vp, sz]),
        "mtx_target_set_sample_offset": (i32, [u64, u32]),
        "mtx_target_get_tuning": (i32, [u64, vp]),
        "mtx_math_eval": (i32, [u32, vp, vp, vp, u64, vp]),
        "mtx_server_create": (i32, [vp, P(u64), C.c_char_p, sz]),
        "mtx_server_submit": (i32, [u64, u64, vp, vp, u32, P(u64), C.c_char_p, sz]),
        "mtx_server_step": (i32, [u64, vp, P(u32), P(u32)]),
        "mtx_server_job_info": (i32, [u64, u64, vp]),
        "mtx_server_job_read_pixels": (i32, [u64, u64, u32, vp, sz, P(sz)]),
        "mtx_server_job_read_accumulators": (i32, [u64, u64, vp, sz, vp, sz]),
        "mtx_server_job_cancel": (i32, [u64, u64]),
        "mtx_server_job_release": (i32, [u64, u64]),
        "mtx_server_stats": (i32, [u64, vp]),
        "mtx_server_graph_stats": (i32, [u64, vp]),
        "mtx_server_destroy": (i32, [u64]),
        "mtx_render_settings_default": (i32, [vp]),
        "mtx_display_settings_default": (i32, [vp]),
        "mtx_jnlm_options_default": (i32, [vp]),
        "mtx_target_compose": (i32, [u64, vp, vp, sz, P(sz)]),
        "mtx_target_denoise_jnlm": (i32, [u64, vp, vp, vp, vp, sz, C.c_char_p, sz]),
        "mtx_context_create": (i32, [P(u64), C.c_char_p, sz]),
        "mtx_context_destroy": (i32, [u64, C.c_char_p, sz]),
        "mtx_context_commit": (i32, [u64, vp, sz, C.c_char_p, sz]),
        "mtx_context_commit_with_settings": (i32, [u64, vp, sz, vp, C.c_char_p, sz]),
        "mtx_context_render": (i32, [u64, u32, C.c_char_p, sz]),
        "mtx_context_reset": (i32, [u64, C.c_char_p, sz]),
        "mtx_context_wait": (i32, [u64, u32, C.c_char_p, sz]),
        "mtx_context_info": (i32, [u64, vp, C.c_char_p, sz]),
        "mtx_context_read_pixels": (i32, [u64, u32, vp, sz, P(sz), C.c_char_p, sz]),
        "mtx_context_submit": (i32, [u64, u32, P(u64), C.c_char_p, sz]),
        "mtx_job_poll": (i32, [u64, u32, C.c_char_p, sz]),
        "mtx_job_info": (i32, [u64, vp, C.c_char_p, sz]),
        "mtx_job_cancel": (i32, [u64, C.c_char_p, sz]),
        "mtx_descriptor_validate": (i32, [vp, sz, vp, C.c_char_p, sz]),
        "mtx_context_inject_failure": (i32, [u64]),
        "mtx_denoiser_create": (i32, [C.c_char_p, u32, P(u64), C.c_char_p, sz]),
        "mtx_denoiser_create_from_weights": (i32, [vp, sz, u32, P(u64), C.c_char_p, sz]),
        "mtx_denoiser_backend": (i32, [u64, P(u32), C.c_char_p, sz]),
        "mtx_denoiser_run": (i32, [u64, u32, u32, vp, vp, vp, vp, sz, C.c_char_p, sz]),
        "mtx_denoiser_forward": (i32, [u64, u32, u32, vp, vp, vp, C.c_char_p, sz]),
        "mtx_denoiser_benchmark": (i32, [u64, u32, u32, u32, P(C.c_double)]),
        "mtx_denoiser_set_tiling": (i32, [u64, u32, u32, u32]),
        "mtx_denoiser_get_tiling": (i32, [u64, u32, u32, u32, P(u32), P(u32)]),
        "mtx_denoiser_set_auxiliary": (i32, [u64, u64, u64, C.c_char_p, sz]),
        "mtx_denoiser_set_tiles": (i32, [u64, u32, u32]),
        "mtx_denoiser_set_input_scale": (i32, [u64, C.c_float]),
        "mtx_denoiser_info": (i32, [u64, P(u32)]),
        "mtx_denoiser_destroy": (i32, [u64, C.c_char_p, sz]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


class DisplaySettings(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("tone_mapping", C.c_uint32), ("exposure", C.c_float),
                ("contrast", C.c_float), ("saturation", C.c_float), ("temperature", C.c_float),
                ("tint", C.c_float), ("vignette", C.c_float), ("gamma22", C.c_uint32)]


class JnlmOptions(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("strength", C.c_float), ("search_radius", C.c_uint32),
                ("patch_radius", C.c_uint32), ("albedo_bandwidth", C.c_float), ("normal_bandwidth", C.c_float),
                ("iterations", C.c_uint32)]
