//! FFI to `libmcrt_hip.so`, the MI355X (gfx950) render hot path of mcrt.
//!
//! `include/mcrt_abi.h` is the authority; this file restates it for Rust.
//! The Pod types are byte-identical to the reference's own `#[repr(C)]`
//! buffer types (`src/scene.rs:22-29`, `src/bvh.rs:54-68`, `src/mesh.rs:24-46`,
//! `src/material.rs:35-63`, `src/camera.rs:35-53`, `src/gpu.rs:164-183`), so a
//! host holding those can pass them with `bytemuck::cast_slice` or a pointer
//! cast.  Every function returns an `i32` status (0 = `MCRT_OK`); the reason
//! of a failure is `mcrt_last_error(ctx)`.  `*_dev` calls take device
//! pointers and a `hipStream_t` as `*mut c_void` (null = the null stream).

#![allow(non_camel_case_types)]

use std::ffi::{c_char, c_int, c_void, CStr};

// ---------------------------------------------------------------------------
// Pod contract types
// ---------------------------------------------------------------------------

/// `src/scene.rs:22-29`: centre.xyz + radius, material index in `mat_and_pad[0]`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct GpuSphere {
    pub center_r: [f32; 4],
    pub mat_and_pad: [u32; 4],
}

/// `src/bvh.rs:54-68`: pre-order flat BVH node, root = 0.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct GpuBvhNode {
    pub aabb_min: [f32; 4],
    pub aabb_max: [f32; 4],
    pub right_or_offset: u32,
    pub prim_count: u32,
    pub _pad: [u32; 2],
}

/// `src/mesh.rs:24-34`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct GpuVertex {
    pub position: [f32; 4],
    pub normal: [f32; 4],
    pub uv: [f32; 4],
}

/// `src/mesh.rs:39-46`; bit 31 of `mat_idx` = back-face cull flag.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct GpuTriangle {
    pub v: [u32; 3],
    pub mat_idx: u32,
}

/// `src/material.rs:35-44`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct GpuMaterial {
    pub type_pad: [u32; 4],
    pub albedo_fuzz: [f32; 4],
    pub ior_pad: [f32; 4],
}

pub const MCRT_MAX_MATERIALS: usize = 64;

/// `src/material.rs:54-63`.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct GpuMaterialData {
    pub mat_count: u32,
    pub n_emissive: u32,
    pub _pad: [u32; 2],
    pub materials: [GpuMaterial; MCRT_MAX_MATERIALS],
}

/// `src/camera.rs:35-53`.
#[repr(C)]
#[derive(Clone, Copy, Debug, Default, PartialEq)]
pub struct CameraUniform {
    pub origin: [f32; 4],
    pub lower_left: [f32; 4],
    pub horizontal: [f32; 4],
    pub vertical: [f32; 4],
    pub defocus_u: [f32; 4],
    pub defocus_v: [f32; 4],
    pub frame_count: u32,
    pub _pad: [u32; 3],
}

/// `src/gpu.rs:164-183`; defaults radius 3, sigma_s 2, sigma_d 0.5, normal_pow 8.
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct DenoiseParams {
    pub radius: i32,
    pub sigma_s: f32,
    pub sigma_d: f32,
    pub normal_pow: f32,
}

impl Default for DenoiseParams {
    fn default() -> Self {
        DenoiseParams { radius: 3, sigma_s: 2.0, sigma_d: 0.5, normal_pow: 8.0 }
    }
}

const _: () = assert!(std::mem::size_of::<GpuSphere>() == 32);
const _: () = assert!(std::mem::size_of::<GpuBvhNode>() == 48);
const _: () = assert!(std::mem::size_of::<GpuVertex>() == 48);
const _: () = assert!(std::mem::size_of::<GpuTriangle>() == 16);
const _: () = assert!(std::mem::size_of::<GpuMaterial>() == 48);
const _: () = assert!(std::mem::size_of::<GpuMaterialData>() == 3088);
const _: () = assert!(std::mem::size_of::<CameraUniform>() == 112);
const _: () = assert!(std::mem::size_of::<DenoiseParams>() == 16);

// ---------------------------------------------------------------------------
// Constants of path_trace.wgsl:36-66 and of the C-ABI
// ---------------------------------------------------------------------------
pub const MCRT_T_MIN: f32 = 1e-4;
pub const MCRT_T_MAX: f32 = 1e9;
pub const MCRT_MAX_FRAME_COUNT: u32 = 65535;
pub const MCRT_MAX_BOUNCES: u32 = 8;
pub const MCRT_MAT_LAMBERTIAN: u32 = 0;
pub const MCRT_MAT_METAL: u32 = 1;
pub const MCRT_MAT_DIELECTRIC: u32 = 2;
pub const MCRT_MAT_EMISSIVE: u32 = 3;
pub const MCRT_BACKFACE_CULL_FLAG: u32 = 0x8000_0000;
pub const MCRT_ENV_MAP_WIDTH: u32 = 1024;
pub const MCRT_ENV_MAP_HEIGHT: u32 = 512;
pub const MCRT_ID_MISS: u32 = 0xFFFF_FFFF;
pub const MCRT_ID_SPHERE_BIT: u32 = 0x8000_0000;

pub const MCRT_OK: c_int = 0;
pub const MCRT_ERR_NO_DEVICE: c_int = -1;
pub const MCRT_ERR_BAD_ARG: c_int = -2;
pub const MCRT_ERR_OOM: c_int = -3;
pub const MCRT_ERR_HIP: c_int = -4;
pub const MCRT_ERR_LIMIT: c_int = -5;
pub const MCRT_ERR_NO_SCENE: c_int = -6;

pub const MCRT_DENOISE_LOSS_LOG_L2: c_int = 0;
pub const MCRT_DENOISE_LOSS_REL_MSE: c_int = 1;
pub const MCRT_DIFF_REPLAY: c_int = 0;
pub const MCRT_DIFF_TAPE: c_int = 1;

/// Opaque context (`mcrt_ctx`).
#[repr(C)]
pub struct McrtCtx {
    _private: [u8; 0],
}

/// Opaque learned denoiser (`mcrt_denoiser`).
#[repr(C)]
pub struct McrtDenoiser {
    _private: [u8; 0],
}

// ---------------------------------------------------------------------------
// Entry points (include/mcrt_abi.h, Part 2)
// ---------------------------------------------------------------------------
#[link(name = "mcrt_hip")]
extern "C" {
    pub fn mcrt_create(out_ctx: *mut *mut McrtCtx, device: c_int) -> c_int;
    pub fn mcrt_destroy(ctx: *mut McrtCtx);
    pub fn mcrt_last_error(ctx: *const McrtCtx) -> *const c_char;
    pub fn mcrt_upload_scene(ctx: *mut McrtCtx, spheres: *const GpuSphere, n_spheres: u32,
                             sphere_orig_index: *const u32, sphere_bvh: *const GpuBvhNode,
                             n_sphere_nodes: u32, verts: *const GpuVertex, n_verts: u32,
                             tris: *const GpuTriangle, n_tris: u32, tri_orig_index: *const u32,
                             mesh_bvh: *const GpuBvhNode, n_mesh_nodes: u32,
                             materials: *const GpuMaterialData, env_rgba32f: *const f32, env_w: u32,
                             env_h: u32, albedo_rgba8: *const u8, n_layers: u32, layer_size: u32) -> c_int;
    pub fn mcrt_render(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32, spp: u32,
                       first_frame: u32, max_depth: u32, accum_in: *const f32, out_accum: *mut f32,
                       out_gbuf: *mut f32) -> c_int;
    pub fn mcrt_render_dev(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32, spp: u32,
                           first_frame: u32, max_depth: u32, accum_in_dev: *const f32, out_accum_dev: *mut f32,
                           out_gbuf_dev: *mut f32, stream: *mut c_void) -> c_int;
    pub fn mcrt_set_nee(ctx: *mut McrtCtx, enabled: c_int) -> c_int;
    pub fn mcrt_render_aux_dev(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32,
                               spp: u32, first_frame: u32, max_depth: u32, accum_in_dev: *const f32,
                               out_accum_dev: *mut f32, out_gbuf_dev: *mut f32, out_albedo_dev: *mut f32,
                               stream: *mut c_void) -> c_int;
    pub fn mcrt_denoise_bilateral(ctx: *mut McrtCtx, params: *const DenoiseParams, width: u32, height: u32,
                                  accum: *const f32, gbuf: *const f32, out: *mut f32) -> c_int;
    pub fn mcrt_denoise_bilateral_dev(ctx: *mut McrtCtx, params: *const DenoiseParams, width: u32,
                                      height: u32, accum_dev: *const f32, gbuf_dev: *const f32,
                                      out_dev: *mut f32, stream: *mut c_void) -> c_int;
    pub fn mcrt_trace_closest(ctx: *mut McrtCtx, origins: *const f32, dirs: *const f32, n: u32, tmin: f32,
                              tmax: f32, out_id: *mut u32, out_t: *mut f32) -> c_int;
    pub fn mcrt_trace_closest_dev(ctx: *mut McrtCtx, origins_dev: *const f32, dirs_dev: *const f32, n: u32,
                                  tmin: f32, tmax: f32, out_id_dev: *mut u32, out_t_dev: *mut f32,
                                  stream: *mut c_void) -> c_int;
    /// No counterpart in the reference: for callers that cannot allocate
    /// (graph capture, serving loops).  Static 32 MiB scratch per GPU;
    /// overlapping calls on one GPU race.
    pub fn mcrt_trace_closest_noalloc_dev(ctx: *mut McrtCtx, origins_dev: *const f32, dirs_dev: *const f32,
                                          n: u32, tmin: f32, tmax: f32, out_id_dev: *mut u32,
                                          out_t_dev: *mut f32, stream: *mut c_void) -> c_int;
    pub fn mcrt_build_mesh_bvh(verts: *const GpuVertex, n_verts: u32, tris: *const GpuTriangle, n_tris: u32,
                               out_nodes: *mut GpuBvhNode, out_n_nodes: *mut u32,
                               out_ordered: *mut GpuTriangle, out_orig_index: *mut u32) -> c_int;
    pub fn mcrt_validate_bvh(nodes: *const GpuBvhNode, n_nodes: u32, n_prims: u32) -> c_int;
    pub fn mcrt_build_sphere_bvh(spheres: *const GpuSphere, n_spheres: u32, out_nodes: *mut GpuBvhNode,
                                 out_n_nodes: *mut u32, out_ordered: *mut GpuSphere,
                                 out_orig_index: *mut u32) -> c_int;
    pub fn mcrt_diff_render(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32,
                            spp: u32, first_frame: u32, max_depth: u32, params: *const f32, n_params: u32,
                            param_of_material: *const i32, out_image: *mut f32) -> c_int;
    pub fn mcrt_diff_render_dev(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32,
                                spp: u32, first_frame: u32, max_depth: u32, params: *const f32, n_params: u32,
                                param_of_material: *const i32, out_image_dev: *mut f32,
                                stream: *mut c_void) -> c_int;
    pub fn mcrt_diff_adjoint(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32,
                             spp: u32, first_frame: u32, max_depth: u32, params: *const f32, n_params: u32,
                             param_of_material: *const i32, d_image: *const f32, out_grad: *mut f32) -> c_int;
    pub fn mcrt_diff_adjoint_dev(ctx: *mut McrtCtx, camera: *const CameraUniform, width: u32, height: u32,
                                 spp: u32, first_frame: u32, max_depth: u32, params: *const f32,
                                 n_params: u32, param_of_material: *const i32, d_image_dev: *const f32,
                                 out_grad_dev: *mut f32, stream: *mut c_void) -> c_int;
    pub fn mcrt_diff_set_mode(ctx: *mut McrtCtx, mode: c_int) -> c_int;
    pub fn mcrt_denoiser_create(ctx: *mut McrtCtx, seed: u32, out: *mut *mut McrtDenoiser) -> c_int;
    pub fn mcrt_denoiser_destroy(dn: *mut McrtDenoiser);
    pub fn mcrt_denoiser_param_count() -> u32;
    pub fn mcrt_denoiser_get_weights(dn: *mut McrtDenoiser, out: *mut f32) -> c_int;
    pub fn mcrt_denoiser_set_weights(dn: *mut McrtDenoiser, weights: *const f32) -> c_int;
    pub fn mcrt_denoiser_infer_dev(dn: *mut McrtDenoiser, batch: u32, width: u32, height: u32,
                                   noisy_dev: *const f32, albedo_dev: *const f32, gbuf_dev: *const f32,
                                   out_dev: *mut f32, stream: *mut c_void) -> c_int;
    pub fn mcrt_denoiser_set_loss(dn: *mut McrtDenoiser, kind: c_int) -> c_int;
    pub fn mcrt_denoiser_train_step_dev(dn: *mut McrtDenoiser, batch: u32, width: u32, height: u32,
                                        noisy_dev: *const f32, albedo_dev: *const f32, gbuf_dev: *const f32,
                                        ref_dev: *const f32, lr: f32, loss_out: *mut f32,
                                        stream: *mut c_void) -> c_int;
    pub fn mcrt_denoiser_gradient(dn: *mut McrtDenoiser, batch: u32, width: u32, height: u32,
                                  noisy_dev: *const f32, albedo_dev: *const f32, gbuf_dev: *const f32,
                                  ref_dev: *const f32, out_grad: *mut f32, loss_out: *mut f32) -> c_int;
    pub fn mcrt_mesh_voxel_chunks(blocks: *const u16, chunks_x: u32, chunks_z: u32, cx0: i32, cz0: i32,
                                  voxel_size: f32, opaque: *const u8, block_material: *const u32,
                                  out_verts: *mut GpuVertex, max_verts: u64, out_tris: *mut GpuTriangle,
                                  max_tris: u64, n_verts: *mut u64, n_tris: *mut u64) -> c_int;
    pub fn mcrt_set_profiling(ctx: *mut McrtCtx, enabled: c_int) -> c_int;
    pub fn mcrt_kernel_stats(ctx: *mut McrtCtx, kernel: *const c_char, total_ms: *mut f64,
                             launches: *mut u64) -> c_int;
    pub fn mcrt_reset_stats(ctx: *mut McrtCtx) -> c_int;
    pub fn mcrt_render_counters(ctx: *mut McrtCtx, segments: *mut u64, samples: *mut u64) -> c_int;
    pub fn mcrt_measure_peaks(device: c_int, copy_bytes: u64, copy_gbs: *mut f64, mfma_bf16_tflops: *mut f64,
                              mfma_bf16_16x16x32_tflops: *mut f64, valu_fp32_tflops: *mut f64) -> c_int;
}

// ---------------------------------------------------------------------------
// Thin safe layer for the headless driver (src/headless.rs:112-386)
// ---------------------------------------------------------------------------

/// A failed call: status code and the library's message.
#[derive(Debug, Clone)]
pub struct McrtError {
    pub code: c_int,
    pub message: String,
}

impl std::fmt::Display for McrtError {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        write!(f, "{} (status {})", self.message, self.code)
    }
}

impl std::error::Error for McrtError {}

fn status(ctx: *const McrtCtx, rc: c_int) -> Result<(), McrtError> {
    if rc == MCRT_OK {
        return Ok(());
    }
    let msg = unsafe { CStr::from_ptr(mcrt_last_error(ctx)) }.to_string_lossy().into_owned();
    Err(McrtError { code: rc, message: msg })
}

/// Owns an `mcrt_ctx`; one per host thread.
pub struct Context {
    raw: *mut McrtCtx,
}

impl Context {
    /// `mcrt_create`: the adapter/device request of `src/headless.rs:135-148`.
    pub fn new(device: i32) -> Result<Self, McrtError> {
        let mut raw = std::ptr::null_mut();
        status(std::ptr::null(), unsafe { mcrt_create(&mut raw, device) })?;
        Ok(Context { raw })
    }

    pub fn raw(&self) -> *mut McrtCtx {
        self.raw
    }

    /// `mcrt_upload_scene` with the reference's BVH outputs (orig indices
    /// unknown: ids are reported in leaf order) and no albedo textures.
    #[allow(clippy::too_many_arguments)]
    pub fn upload_scene(&mut self, spheres: &[GpuSphere], sphere_bvh: &[GpuBvhNode], verts: &[GpuVertex],
                        tris: &[GpuTriangle], mesh_bvh: &[GpuBvhNode], materials: &GpuMaterialData,
                        env_rgba32f: &[f32], env_w: u32, env_h: u32) -> Result<(), McrtError> {
        if env_rgba32f.len() < (env_w as usize) * (env_h as usize) * 4 {
            return Err(McrtError { code: MCRT_ERR_BAD_ARG, message: "environment map too small".into() });
        }
        status(self.raw, unsafe {
            mcrt_upload_scene(self.raw, spheres.as_ptr(), spheres.len() as u32, std::ptr::null(),
                              sphere_bvh.as_ptr(), sphere_bvh.len() as u32, verts.as_ptr(), verts.len() as u32,
                              tris.as_ptr(), tris.len() as u32, std::ptr::null(), mesh_bvh.as_ptr(),
                              mesh_bvh.len() as u32, materials, env_rgba32f.as_ptr(), env_w, env_h,
                              std::ptr::null(), 0, 0)
        })
    }

    /// `mcrt_render`: the spp loop and readback of `src/headless.rs:249-348`.
    /// Returns the RGBA32F running mean (and the G-buffer if asked).
    pub fn render(&mut self, camera: &CameraUniform, width: u32, height: u32, spp: u32, first_frame: u32,
                  max_depth: u32, accum_in: Option<&[f32]>, want_gbuf: bool)
                  -> Result<(Vec<f32>, Option<Vec<f32>>), McrtError> {
        let n = (width as usize) * (height as usize) * 4;
        if let Some(a) = accum_in {
            if a.len() < n {
                return Err(McrtError { code: MCRT_ERR_BAD_ARG, message: "accum_in too small".into() });
            }
        }
        let mut accum = vec![0f32; n];
        let mut gbuf = if want_gbuf { Some(vec![0f32; n]) } else { None };
        let gptr = gbuf.as_mut().map_or(std::ptr::null_mut(), |g| g.as_mut_ptr());
        status(self.raw, unsafe {
            mcrt_render(self.raw, camera, width, height, spp, first_frame, max_depth,
                        accum_in.map_or(std::ptr::null(), |a| a.as_ptr()), accum.as_mut_ptr(), gptr)
        })?;
        Ok((accum, gbuf))
    }

    /// `mcrt_denoise_bilateral`: `cs_denoise` of `src/gpu.rs:1150-1169`.
    pub fn denoise_bilateral(&mut self, params: &DenoiseParams, width: u32, height: u32, accum: &[f32],
                             gbuf: &[f32]) -> Result<Vec<f32>, McrtError> {
        let n = (width as usize) * (height as usize) * 4;
        if accum.len() < n || gbuf.len() < n {
            return Err(McrtError { code: MCRT_ERR_BAD_ARG, message: "image buffers too small".into() });
        }
        let mut out = vec![0f32; n];
        status(self.raw, unsafe {
            mcrt_denoise_bilateral(self.raw, params, width, height, accum.as_ptr(), gbuf.as_ptr(),
                                   out.as_mut_ptr())
        })?;
        Ok(out)
    }

    /// `mcrt_trace_closest`: `world_hit` for a batch of rays.
    pub fn trace_closest(&mut self, origins: &[f32], dirs: &[f32], tmin: f32, tmax: f32)
                         -> Result<(Vec<u32>, Vec<f32>), McrtError> {
        let n = origins.len() / 3;
        if dirs.len() < 3 * n {
            return Err(McrtError { code: MCRT_ERR_BAD_ARG, message: "dirs shorter than origins".into() });
        }
        let mut ids = vec![0u32; n];
        let mut ts = vec![0f32; n];
        status(self.raw, unsafe {
            mcrt_trace_closest(self.raw, origins.as_ptr(), dirs.as_ptr(), n as u32, tmin, tmax,
                               ids.as_mut_ptr(), ts.as_mut_ptr())
        })?;
        Ok((ids, ts))
    }

    /// `mcrt_set_nee`: next-event estimation for later renders (default off).
    pub fn set_nee(&mut self, enabled: bool) -> Result<(), McrtError> {
        status(self.raw, unsafe { mcrt_set_nee(self.raw, enabled as c_int) })
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        unsafe { mcrt_destroy(self.raw) };
    }
}
