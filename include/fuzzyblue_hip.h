/*
 * fuzzyblue_hip.h — C ABI of the MI355X-native precomputed atmospheric
 * scattering engine (Bruneton–Neyret LUTs on gfx950).
 *
 * The reference library (Ralith/fuzzyblue) exposes a Rust API over Vulkan:
 * Builder, Atmosphere, PendingAtmosphere, Parameters, Renderer and
 * DrawParameters (src/lib.rs:8-12).  Every entry point below names the
 * reference interface it replaces (file:line in the reference tree).
 *
 * Conventions
 *  - Distances are kilometres, as in the reference (src/precompute.rs:680).
 *  - Every call returns an fb_status; on failure a thread-local message is
 *    available from fb_last_error().  Nothing aborts the process.
 *  - Work-issuing calls only ENQUEUE on the caller's stream (a hipStream_t
 *    passed as fb_stream; NULL means the device's default stream), mirroring
 *    the reference's record-only Atmosphere::build (src/precompute.rs:1077).
 *    No call synchronises the device except the download / wait helpers.
 *  - All LUTs live in HBM as row-major float4 (RGBA32F) arrays (RGBA
 *    binary16 for the scattering-sized LUTs with FB_FLAG_F16_STORAGE): x
 *    fastest, then y, then z.  The scattering LUT is the reference's 4D table packed
 *    as a 3D extent (NU*MUS) x MU x R with x = nu*MUS + mu_s
 *    (src/precompute.rs:786-791, shaders/scattering.h:120-131), identical to
 *    the linear readback layout of examples/dump.rs:175-193.
 *  - Handles are library-owned.  Device memory belongs to the atmosphere or
 *    pending object that allocated it.
 */
#ifndef FUZZYBLUE_HIP_H_
#define FUZZYBLUE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FB_ABI_VERSION 1u

/* Status codes.  The reference surfaces no errors (every Vulkan failure is an
 * unwrap() panic, e.g. src/precompute.rs:590-598); the C ABI reports them. */
typedef enum fb_status {
    FB_OK = 0,
    FB_EINVAL = 1,     /* invalid argument or parameter set */
    FB_ENOMEM = 2,     /* device or host allocation failed */
    FB_EBUDGET = 3,    /* request does not fit the memory budget */
    FB_EHIP = 4,       /* a HIP runtime call failed */
    FB_EBUSY = 5,      /* object still in use by in-flight GPU work */
    FB_ENOTREADY = 6,  /* work not complete yet (non-blocking query) */
    FB_ENOENT = 7,     /* unknown handle / id */
    FB_ENODEV = 8      /* no usable HIP device */
} fb_status;

typedef void *fb_stream; /* a hipStream_t; NULL = default stream */
typedef void *fb_event;  /* a hipEvent_t */

/* ---------------------------------------------------------------------- */
/* Parameters — src/precompute.rs:655-769 (Parameters, DensityProfile,      */
/* DensityProfileLayer).  Same fields in declaration order, minus the four  */
/* Vulkan synchronisation fields (usage, dst_stage_mask, dst_access_mask,   */
/* layout, :691-698) which have no meaning on HIP streams.  The reference   */
/* spells two fields "absorbtion_*"; C uses "absorption_*".                  */
/* ---------------------------------------------------------------------- */

/* Density clamp(exp_term*exp(exp_scale*h) + linear_term*h + constant_term,
 * 0, 1) at altitude h (km) — src/precompute.rs:655-665. */
typedef struct fb_density_profile_layer {
    float width;
    float exp_term;
    float exp_scale;
    float linear_term;
    float constant_term;
} fb_density_profile_layer;

/* Two layers bottom to top; layer 0 applies below layers[0].width
 * (src/precompute.rs:667-676, shaders/params.h:95-99). */
typedef struct fb_density_profile {
    fb_density_profile_layer layers[2];
} fb_density_profile;

/* fb_parameters.flags */
#define FB_FLAG_PRECISE 0x1u /* FP64 ray-sphere geometry and horizon
                                parameterisation, compensated sums */
/* The reference's storage formats (src/precompute.rs:1215): the scattering
 * LUT and the four 3D scratch LUTs (delta Rayleigh, delta Mie, delta
 * multiple scattering, scattering density) are RGBA binary16, 8 bytes per
 * texel, every store rounded to nearest even as the reference's
 * R16G16B16A16_SFLOAT image stores are; arithmetic stays FP32 (FP64
 * geometry with FB_FLAG_PRECISE).  Transmittance and the irradiance LUTs
 * stay RGBA FP32, as in the reference (:1170, :1191).  Views of the
 * binary16 LUTs report texel_bytes = 8. */
#define FB_FLAG_F16_STORAGE 0x2u

typedef struct fb_parameters {
    uint32_t order; /* number of light bounces; 1 = single scattering only */

    uint32_t transmittance_mu_size;
    uint32_t transmittance_r_size;
    uint32_t scattering_r_size;
    uint32_t scattering_mu_size;
    uint32_t scattering_mu_s_size;
    uint32_t scattering_nu_size;
    uint32_t irradiance_mu_s_size;
    uint32_t irradiance_r_size;

    float solar_irradiance[3];
    float sun_angular_radius; /* radians, < 0.1 */
    float bottom_radius;      /* km */
    float top_radius;         /* km */
    fb_density_profile rayleigh_density;
    float rayleigh_scattering[3]; /* 1/km */
    fb_density_profile mie_density;
    float mie_scattering[3]; /* 1/km */
    float mie_extinction[3]; /* 1/km */
    float mie_phase_function_g;
    fb_density_profile absorption_density;
    float absorption_extinction[3]; /* 1/km */
    float ground_albedo[3];
    float mu_s_min;

    uint32_t flags; /* FB_FLAG_* */
} fb_parameters;

/* Parameters::default — src/precompute.rs:849-935 (Earth, km units, order 4,
 * LUTs 256x64 / 64x16 / 32x128x32x8). */
void fb_parameters_default(fb_parameters *out);

/* Range checks the reference leaves to the caller: every size >= 2,
 * scattering_mu_size >= 4, 0 < bottom < top, order >= 1, |g| < 1,
 * -1 <= mu_s_min < 0 ... (the reference divides by size-1, MU/2-1 and
 * mu_s_min at shaders/util.h:37, scattering.h:36,54). */
fb_status fb_parameters_validate(const fb_parameters *p);

/* Parameters::{transmittance,irradiance,scattering}_extent —
 * src/precompute.rs:771-793.  Any output pointer may be NULL. */
void fb_parameters_extents(const fb_parameters *p, uint32_t transmittance[2],
                           uint32_t irradiance[2], uint32_t scattering[3]);

/* Bytes of HBM a LUT set occupies: resident outputs (transmittance,
 * irradiance, scattering) and build-time scratch (five deltas and the
 * multiple-scattering row table).  Returns resident + scratch. */
size_t fb_parameters_footprint(const fb_parameters *p, size_t *resident,
                               size_t *scratch);

/* ---------------------------------------------------------------------- */
/* Builder — src/precompute.rs:15-30, Builder::new :61-68.                  */
/* ---------------------------------------------------------------------- */
typedef struct fb_builder fb_builder;
typedef struct fb_atmosphere fb_atmosphere;
typedef struct fb_pending_atmosphere fb_pending_atmosphere;

/* Builder::new(instance, device, cache, physical, gfx_qf, compute_qf)
 * (src/precompute.rs:61-68).  Binds a HIP device and owns the kernel
 * handles, the LUT allocator and the tile-shape table. */
fb_status fb_builder_create(int hip_device, fb_builder **out);
void fb_builder_destroy(fb_builder *b);

/* ---------------------------------------------------------------------- */
/* Precompute — Atmosphere::build(builder, cmd, &Parameters)                */
/* -> PendingAtmosphere (src/precompute.rs:1077-1081).  Enqueue only.       */
/* ---------------------------------------------------------------------- */
fb_status fb_atmosphere_build(fb_builder *b, fb_stream stream,
                              const fb_parameters *params,
                              fb_pending_atmosphere **out);

/* Many atmospheres of uneven resolution and order in one launch per pass
 * (no reference counterpart: the reference builds one atmosphere per
 * recorded command buffer).  out[i] receives the pending handle of
 * params[i]; all share one completion event. */
fb_status fb_atmosphere_build_batch(fb_builder *b, fb_stream stream,
                                    const fb_parameters *params, size_t n,
                                    fb_pending_atmosphere **out);

/* The precompute of one atmosphere into caller-owned output LUTs, with no
 * device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, serving
 * loops and harnesses that forbid hipMalloc).  The caller passes the three
 * output LUTs (transmittance and irradiance RGBA32F, scattering RGBA32F or,
 * with FB_FLAG_F16_STORAGE, RGBA binary16, in the layouts above) with their
 * sizes in bytes; every transient LUT and table lives in a static device
 * array of fb_build_into_scratch_limit() bytes (96 MiB: the default LUT sizes
 * fit, larger sets return FB_EINVAL), one per GPU, held from the first
 * fb_builder_create on that GPU until the process ends.  Enqueue only; no
 * event or object is created, so the caller orders later work with its own
 * stream.  Calls that overlap on one GPU share the array and race: enqueue
 * them on one stream.  LUTs are bit-identical to fb_atmosphere_build's. */
fb_status fb_atmosphere_build_into(fb_builder *b, fb_stream stream,
                                   const fb_parameters *params,
                                   void *transmittance, size_t transmittance_bytes,
                                   void *irradiance, size_t irradiance_bytes,
                                   void *scattering, size_t scattering_bytes);
size_t fb_build_into_scratch_limit(void);

/* Re-run the recorded precompute of a pending atmosphere into the same
 * buffers (the reference's bench re-submits one recorded command buffer:
 * benches/precompute.rs:138-148).  Uses the captured hipGraph. */
fb_status fb_pending_relaunch(fb_pending_atmosphere *pa, fb_stream stream);

/* Block sizes of the tunable passes for one LUT-size class (the precision
 * flag and the eight sizes of fb_parameters).  No reference counterpart:
 * the reference's shaders fix their workgroup sizes (8x8, 4x4x4).  Every
 * choice gives bit-identical LUTs; builds (plain and batched) started
 * after the tuning is set use it. */
typedef struct fb_tuning {
    uint32_t single_scattering_threads;   /* 64, 128 or 256 per (r, mu) row */
    uint32_t density_order2_threads;      /* 64, 128 or 256 per (r, mu_s) tile */
    uint32_t density_threads;             /* the same for orders >= 3 */
    uint32_t indirect_order1_waves;       /* 1, 2 or 4 wavefronts per block */
    uint32_t indirect_waves;              /* the same for orders >= 2 */
    uint32_t multiple_scattering_threads; /* 64 .. 1024 (a multiple of 64) per
                                             (r, mu) row: at most 4 texels per
                                             thread, and above 256 at most one
                                             wavefront more than the row */
    uint32_t xcd_remap; /* bit p set (p = FB_PASS_SINGLE_SCATTERING,
                           FB_PASS_SCATTERING_DENSITY or
                           FB_PASS_MULTIPLE_SCATTERING): that pass gives each
                           of the 8 XCDs a contiguous share of its blocks */
    float best_ms[6]; /* fb_builder_autotune: measured ms per launch of each
                         chosen shape, in field order; 0 otherwise */
} fb_tuning;

/* Time every candidate block size of each tunable pass on a scratch LUT
 * set of p's sizes (enqueued on `stream`, synchronised), keep the fastest
 * in the builder's table, and return the choice in *out (may be NULL). */
fb_status fb_builder_autotune(fb_builder *b, fb_stream stream,
                              const fb_parameters *p, fb_tuning *out);
/* The block sizes builds of p's size class use; FB_ENOENT (with the
 * defaults in *out) when the class was never tuned. */
fb_status fb_builder_get_tuning(fb_builder *b, const fb_parameters *p,
                                fb_tuning *out);
/* Install block sizes (0 = default) for p's size class, e.g. a table
 * saved from an earlier autotune; FB_EINVAL if one is not valid there. */
fb_status fb_builder_set_tuning(fb_builder *b, const fb_parameters *p,
                                const fb_tuning *t);

/* Pass identifiers of the precompute schedule (src/precompute.rs:1726-1999). */
typedef enum fb_pass {
    FB_PASS_TRANSMITTANCE = 0,       /* shaders/transmittance.comp */
    FB_PASS_DIRECT_IRRADIANCE = 1,   /* shaders/direct_irradiance.comp */
    FB_PASS_SINGLE_SCATTERING = 2,   /* shaders/single_scattering.comp */
    FB_PASS_CLEAR_IRRADIANCE = 3,    /* src/precompute.rs:1802-1816 */
    FB_PASS_SCATTERING_DENSITY = 4,  /* shaders/scattering_density.comp */
    FB_PASS_INDIRECT_IRRADIANCE = 5, /* shaders/indirect_irradiance.comp */
    FB_PASS_MULTIPLE_SCATTERING = 6, /* shaders/multiple_scattering.comp */
    FB_PASS_GEOMETRY = 7             /* no reference counterpart: per-set
                                        geometry tables built once after the
                                        transmittance pass: the density
                                        pass's taps and ground terms, and the
                                        51 samples of every (r, mu) row that
                                        single and multiple scattering march */
} fb_pass;

typedef struct fb_pass_timing {
    uint32_t pass;  /* fb_pass */
    uint32_t order; /* scattering order of the launch (1 for the first passes) */
    float ms;       /* GPU time between the events around the launch */
} fb_pass_timing;

/* Tracing: re-run the recorded schedule into the same buffers with HIP
 * events between launches, wait, and report each launch's duration in
 * schedule order (the reference has no timing; its harness brackets a
 * whole submit with device_wait_idle, benches/precompute.rs:138-148). */
fb_status fb_pending_profile(fb_pending_atmosphere *pa, fb_stream stream,
                             fb_pass_timing *out, size_t cap, size_t *n);

/* PendingAtmosphere::atmosphere (src/precompute.rs:2204): the atmosphere
 * that will be valid once the enqueued work completes. */
const fb_atmosphere *fb_pending_peek(const fb_pending_atmosphere *pa);

/* FB_OK when complete, FB_ENOTREADY otherwise (non-blocking). */
fb_status fb_pending_query(const fb_pending_atmosphere *pa);

/* Block the host until the precompute completed. */
fb_status fb_pending_wait(const fb_pending_atmosphere *pa);

/* Record the completion event on another stream's timeline: stream waits
 * for the precompute (replaces PendingAtmosphere::acquire_ownership's
 * queue-family transfer, src/precompute.rs:2142-2201). */
fb_status fb_pending_stream_wait(const fb_pending_atmosphere *pa,
                                 fb_stream consumer);

/* The same hand-over as an event the caller owns (acquire_ownership,
 * src/precompute.rs:2147-2152, for callers that synchronise on events):
 * event completes once the precompute has completed.  The library records
 * it on a private stream after a wait for the precompute, so the caller's
 * streams are not touched.  Re-recording it later moves it as usual. */
fb_status fb_pending_record_ready_event(const fb_pending_atmosphere *pa,
                                        fb_event event);

/* PendingAtmosphere::assert_ready (src/precompute.rs:2209): consumes the
 * pending object, returns the atmosphere and releases the scratch LUTs.
 * Returns FB_EBUSY (and leaves the pending object intact) if the work has
 * not completed. */
fb_status fb_pending_assert_ready(fb_pending_atmosphere *pa,
                                  fb_atmosphere **out);

/* Drop for PendingAtmosphere (src/precompute.rs:2122-2140).  Waits for the
 * enqueued work before freeing, so it never frees memory in use. */
void fb_pending_destroy(fb_pending_atmosphere *pa);

/* LUT identifiers.  The first three are the atmosphere's outputs
 * (Atmosphere::{transmittance,irradiance,scattering}, src/precompute.rs:
 * 2075-2101); the rest are the pending object's scratch textures
 * (src/precompute.rs:2111-2120), readable until assert_ready/destroy. */
typedef enum fb_lut {
    FB_LUT_TRANSMITTANCE = 0,
    FB_LUT_IRRADIANCE = 1,
    FB_LUT_SCATTERING = 2,
    FB_LUT_DELTA_IRRADIANCE = 3,
    FB_LUT_DELTA_RAYLEIGH = 4,
    FB_LUT_DELTA_MIE = 5,
    FB_LUT_DELTA_MULTIPLE_SCATTERING = 6,
    FB_LUT_SCATTERING_DENSITY = 7
} fb_lut;

typedef struct fb_lut_view {
    void *device_ptr;  /* RGBA texels, row-major x,y,z */
    uint32_t width;
    uint32_t height;
    uint32_t depth;    /* 1 for 2D LUTs */
    uint32_t texel_bytes; /* 16 (float32 x4), or 8 (binary16 x4) for the
                             scattering-sized LUTs with FB_FLAG_F16_STORAGE */
    size_t bytes;
} fb_lut_view;

/* Any LUT of a pending atmosphere, scratch included (for inspection once
 * its work completed).  FB_LUT_DELTA_IRRADIANCE is the last bounce's
 * delta irradiance (the direct irradiance for order 1). */
fb_status fb_pending_lut(const fb_pending_atmosphere *pa, fb_lut which,
                         fb_lut_view *out);

/* Atmosphere accessors (src/precompute.rs:2075-2101). */
fb_status fb_atmosphere_lut(const fb_atmosphere *a, fb_lut which,
                            fb_lut_view *out);
fb_status fb_atmosphere_parameters(const fb_atmosphere *a, fb_parameters *out);

/* Copy a LUT to host memory in the examples/dump.rs:110-126 linear layout.
 * format 0: float32 x4 (16 B/texel); format 1: float16 x4 (8 B/texel, the
 * reference's R16G16B16A16_SFLOAT scattering format, src/precompute.rs:1215).
 * A LUT stored in the other format is converted (binary16 widens exactly,
 * float32 rounds to nearest even).  Synchronises with the atmosphere's
 * completion event. */
#define FB_FORMAT_F32 0u
#define FB_FORMAT_F16 1u
fb_status fb_atmosphere_download(const fb_atmosphere *a, fb_lut which,
                                 uint32_t format, void *host, size_t bytes);

/* LUT files: the images examples/dump.rs:331-367 writes with the openexr
 * crate, one per LUT — width x height, channels "R","G","B","A" for a 2D
 * LUT and "{z}.R" .. "{z}.A" for depth slice z of the scattering LUT, the
 * linear layout of examples/dump.rs:175-193 — as uncompressed single-part
 * scanline OpenEXR files.  Transmittance and irradiance are 32-bit float;
 * the scattering LUT is 16-bit half with FB_FORMAT_F16 (the dump's format)
 * or 32-bit float with FB_FORMAT_F32.  Two header attributes,
 * "fuzzyblueLut" and "fuzzyblueParameters" (the fb_parameters bytes, host
 * byte order), let fb_atmosphere_load rebuild the set.  Save synchronises
 * with the atmosphere's completion event. */
fb_status fb_atmosphere_save(const fb_atmosphere *a, fb_lut which,
                             uint32_t format, const char *path);
/* A finished atmosphere from three files written by fb_atmosphere_save
 * (same atmosphere, any scattering format): the LUTs are uploaded into the
 * builder's pool in the saved parameters' storage format (binary16 widens
 * exactly, float32 rounds to nearest even) and the set is ready on return.
 * FB_EINVAL for files of another layout, size or atmosphere. */
fb_status fb_atmosphere_load(fb_builder *b, const char *transmittance_path,
                             const char *irradiance_path,
                             const char *scattering_path, fb_atmosphere **out);

/* A LUT shared with another API as a Linux dma-buf descriptor, in place of
 * the reference's images that the application samples itself
 * (src/precompute.rs:1167-1228: created with the caller's extra usage
 * flags, Parameters::usage, and handed out by Atmosphere::{transmittance,
 * irradiance,scattering}, src/precompute.rs:2075-2101).  The LUT is copied
 * (device to device, ordered after the atmosphere's completion on `stream`;
 * the call returns when the copy is done) into an allocation of its own made
 * exportable as a POSIX file descriptor: bytes from offset, texels in the
 * linear x, y, z order with the pitches given, in vk_format (97
 * R16G16B16A16_SFLOAT, 109 R32G32B32A32_SFLOAT), zero padding up to
 * alloc_bytes.  Import fd with VK_EXT_external_memory_dma_buf
 * (VkImportMemoryFdInfoKHR, allocationSize alloc_bytes; Vulkan takes
 * ownership of fd on success) or hipImportExternalMemory.  device_ptr is the
 * library's own mapping of the same memory.  The export does not depend on
 * the atmosphere afterwards. */
typedef struct fb_lut_export {
    int fd;               /* owned by the caller until passed on */
    uint32_t width;
    uint32_t height;
    uint32_t depth;       /* 1 for 2D LUTs */
    uint32_t texel_bytes; /* 16, or 8 for binary16 storage */
    uint32_t vk_format;
    size_t bytes;         /* texels * texel_bytes */
    size_t alloc_bytes;   /* size of the shared allocation */
    size_t offset;        /* of the first texel in the shared allocation */
    size_t row_pitch;
    size_t slice_pitch;
    void *device_ptr;
    void *opaque;         /* the library's handle, for fb_lut_export_release */
} fb_lut_export;
fb_status fb_atmosphere_export(const fb_atmosphere *a, fb_lut which,
                               fb_stream stream, fb_lut_export *out);
/* Closes fd if it is still >= 0 (set it to -1 once passed on) and drops the
 * library's mapping and handle; the memory lives until every importer has
 * released it. */
void fb_lut_export_release(fb_lut_export *e);

/* Drop for Atmosphere (src/precompute.rs:1059-1073). */
void fb_atmosphere_destroy(fb_atmosphere *a);

/* ---------------------------------------------------------------------- */
/* Rendering — Renderer::draw (src/render.rs:209-236) running              */
/* shaders/render_sky.frag:24-35 per pixel.                                  */
/* ---------------------------------------------------------------------- */

/* DrawParameters (src/render.rs:243-252).  inverse_viewproj is column-major
 * like the GLSL mat4 push constant: inverse_viewproj[column][row]. */
typedef struct fb_draw_parameters {
    float inverse_viewproj[4][4];
    float camera_position[3]; /* km, planet frame */
    float sun_direction[3];   /* unit vector */
} fb_draw_parameters;

/* fb_render_sky flags */
#define FB_RENDER_COMPOSITE 0x1u /* color = radiance + color * transmittance,
                                    the reference's dual-source blend
                                    (src/render.rs:124-136) */

/* One frame of render_sky.frag over a width x height target.  depth: device
 * float per pixel (the subpassInput depth, render_sky.frag:26), or NULL for
 * an infinitely distant background (reverse-Z depth 0).  color and
 * transmittance: device float4 per pixel, row-major, row 0 = top (Vulkan
 * framebuffer origin).  transmittance may be NULL when compositing. */
fb_status fb_render_sky(fb_builder *b, fb_stream stream, const fb_atmosphere *a,
                        const fb_draw_parameters *draw, const float *depth,
                        uint32_t width, uint32_t height, float *color,
                        float *transmittance, uint32_t flags);

/* Per-ray sky lookup, GetSkyRadiance (shaders/render_sky.h:45-109) or, when
 * points != NULL, GetSkyRadianceToPoint (:111-191).  All arrays are device
 * float3 (packed, 12 B) except the float4 outputs.  cameras/suns may hold
 * one entry (stride 0) or n entries, as given by the *_stride arguments
 * (0 or 1). */
fb_status fb_render_sky_rays(fb_builder *b, fb_stream stream,
                             const fb_atmosphere *a, const float *cameras,
                             uint32_t camera_stride, const float *views,
                             const float *points, const float *suns,
                             uint32_t sun_stride, size_t n, float *radiance,
                             float *transmittance);

/* fb_render_sky_rays with light shafts: the shadow_length term the
 * reference keeps commented out (shaders/render_sky.h:81-106, :160,
 * :173-177; Bruneton & Neyret 2008, Eqs. 17-18).  shadow_lengths (device
 * float, km, one per ray; NULL = all 0) is the length l of the ray's
 * shadowed part.  GetSkyRadiance omits the scattering of the first l km
 * from the viewer (the lookup moves to x + l v and is attenuated by
 * T(x, x + l v)); GetSkyRadianceToPoint omits the last l km before the
 * point (its second lookup moves to d - l, attenuated by T(x, x_{d-l})); the
 * transmittance output is unchanged.  l = 0 gives fb_render_sky_rays bit for
 * bit. */
fb_status fb_render_sky_rays_shadowed(fb_builder *b, fb_stream stream,
                                      const fb_atmosphere *a, const float *cameras,
                                      uint32_t camera_stride, const float *views,
                                      const float *points, const float *suns,
                                      uint32_t sun_stride, const float *shadow_lengths,
                                      size_t n, float *radiance, float *transmittance);

/* Synthetic frames for throughput measurement: frame f has camera
 * (0, 0, bottom + altitude[f]) and sun (sin zs[f], 0, cos zs[f]); pixel i of
 * frame f looks along a direction uniform on the sphere drawn from a
 * counter-based hash of (seed, f, i) (fb_ray_direction below).  Writes
 * n_frames*pixels float4 radiance and transmittance (GetSkyRadiance).  When
 * the LUTs' per-frame slice fits a block's LDS (default sizes), the call
 * takes n_frames * (at most 64 KB) of stream-ordered scratch on `stream`
 * (hipMallocAsync, freed after its launches); FB_ENOMEM if that fails. */
fb_status fb_render_sky_frames(fb_builder *b, fb_stream stream,
                               const fb_atmosphere *a,
                               const float *altitudes_km,
                               const float *sun_zenith_rad, uint32_t n_frames,
                               uint32_t pixels, uint64_t seed, float *radiance,
                               float *transmittance);

/* The direction generator fb_render_sky_frames uses, on the host (so a
 * test can rebuild the same rays). */
void fb_ray_direction(uint64_t seed, uint32_t frame, uint32_t pixel,
                      float out[3]);

/* GetSunAndSkyIrradiance (shaders/render_lighting.h:10-28) for n surface
 * points (device float3 arrays; suns stride 0 or 1).  Outputs device float4:
 * direct sun irradiance and sky irradiance. */
fb_status fb_sun_and_sky_irradiance(fb_builder *b, fb_stream stream,
                                    const fb_atmosphere *a, const float *points,
                                    const float *normals, const float *suns,
                                    uint32_t sun_stride, size_t n,
                                    float *sun_irradiance,
                                    float *sky_irradiance);

/* ---------------------------------------------------------------------- */
/* Serving: continuous batching of precompute and render requests with     */
/* HBM-budget admission and LRU eviction of resident LUT sets.  No          */
/* reference counterpart (the reference has no request loop).               */
/* ---------------------------------------------------------------------- */
typedef struct fb_server fb_server;

typedef struct fb_server_config {
    size_t hbm_budget_bytes;   /* resident LUT sets + in-flight scratch and
                                  render outputs */
    uint32_t max_inflight;     /* precompute slots on the device */
    uint32_t n_streams;        /* device streams the server schedules on */
    uint32_t keep_outputs;     /* nonzero: keep each render's radiance until
                                  fb_server_render_result fetches it */
    uint32_t precompute_share_pct; /* 1..99: every precompute launch is
                                  persistent and holds at most this share of
                                  the device's resident-block slots, so renders
                                  find free slots on every CU; 0 or >= 100: one
                                  block per tile (full occupancy) */
    uint32_t admit_min_free;   /* 0 or 1: admit precomputes as soon as a slot
                                  frees (least queueing); k > 1: while
                                  precomputes are in flight, admit the next
                                  group only once k slots are free, so a bulk
                                  submission runs as batches of at least k
                                  instead of one small batch per completion
                                  (an idle server admits at once) */
} fb_server_config;

typedef enum fb_request_kind {
    FB_REQUEST_PRECOMPUTE = 0,
    FB_REQUEST_RENDER = 1
} fb_request_kind;

typedef struct fb_request_info {
    uint64_t id;
    uint32_t kind;       /* fb_request_kind */
    int32_t status;      /* fb_status of the finished request */
    uint64_t lut_id;     /* LUT set produced (precompute) or read (render) */
    double t_submit;     /* seconds, host monotonic clock */
    double t_start;      /* admitted onto the device */
    double t_done;       /* completion observed by the server */
    float gpu_ms;        /* render: GPU time from the launch reaching the
                            render stream to the kernel's end (queueing for
                            CUs included); precompute: 0 */
    float gpu_wait_ms;   /* render: GPU time from admission (a timestamp on an
                            idle stream) to the launch reaching the render
                            stream; precompute: 0 */
} fb_request_info;

fb_status fb_server_create(fb_builder *b, const fb_server_config *cfg,
                           fb_server **out);
void fb_server_destroy(fb_server *s);
/* Submit a precompute; *lut_id receives the id the LUT set will have. */
fb_status fb_server_submit_precompute(fb_server *s, const fb_parameters *p,
                                      uint64_t *request_id, uint64_t *lut_id);
/* Submit n precomputes at once (no counterpart in the reference, which
 * has no request loop): the loop sees them together and admits as many as
 * its in-flight cap and memory budget allow as one batch, where one call
 * per request lets it start on the first ones alone.  All parameter sets
 * are validated first; on an invalid one nothing is enqueued.  Ids are
 * written to request_ids[i] / lut_ids[i] (either may be NULL). */
fb_status fb_server_submit_precomputes(fb_server *s, const fb_parameters *ps,
                                       size_t n, uint64_t *request_ids,
                                       uint64_t *lut_ids);
/* Submit a sky render of width x height synthetic rays (as
 * fb_render_sky_frames, one frame) against resident LUT set lut_id. */
fb_status fb_server_submit_render(fb_server *s, uint64_t lut_id,
                                  float altitude_km, float sun_zenith_rad,
                                  uint32_t width, uint32_t height,
                                  uint64_t seed, uint64_t *request_id);
/* Move up to cap finished requests into out; returns the count in *n. */
fb_status fb_server_poll(fb_server *s, fb_request_info *out, size_t cap,
                         size_t *n);
/* Block until every submitted request finished. */
fb_status fb_server_drain(fb_server *s);
/* Read back a resident LUT set (for the alone-vs-mix checks). */
fb_status fb_server_download(fb_server *s, uint64_t lut_id, fb_lut which,
                             void *host, size_t bytes);
/* Read back a finished render's radiance (float4 per pixel). */
fb_status fb_server_render_result(fb_server *s, uint64_t request_id,
                                  void *host, size_t bytes);
typedef struct fb_server_stats {
    uint64_t submitted, completed, rejected, evictions;
    size_t hbm_in_use, hbm_peak;
    uint32_t resident_sets;
    /* Longest single step of the worker loop, in ms: retiring finished work,
     * admitting renders, admitting precomputes (host time, HIP calls
     * included); and how many steps of any kind took over 2 ms. */
    float step_max_ms[3];
    uint64_t slow_steps;
} fb_server_stats;
fb_status fb_server_get_stats(fb_server *s, fb_server_stats *out);
/* Ids of the LUT sets currently resident (finished and not evicted). */
fb_status fb_server_resident(fb_server *s, uint64_t *ids, size_t cap,
                             size_t *n);

/* ---------------------------------------------------------------------- */
/* Device plumbing for callers without their own HIP bindings (tests).     */
/* ---------------------------------------------------------------------- */
fb_status fb_device_count(int *out);
fb_status fb_device_alloc(size_t bytes, void **out);
fb_status fb_device_free(void *ptr);
fb_status fb_memcpy_h2d(void *dst, const void *src, size_t bytes);
fb_status fb_memcpy_d2h(void *dst, const void *src, size_t bytes);
fb_status fb_memset_d(void *dst, int value, size_t bytes);
/* Device-to-device copy enqueued on a stream (also the known-size kernel
 * that calibrates HBM byte counters). */
fb_status fb_device_copy(void *dst, const void *src, size_t bytes,
                         fb_stream stream);
fb_status fb_stream_create(fb_stream *out);
fb_status fb_stream_destroy(fb_stream s);
fb_status fb_stream_synchronize(fb_stream s);
fb_status fb_device_synchronize(void);

/* Timed event pair helpers (for measurement on the library's streams). */
fb_status fb_event_create(fb_event *out);
fb_status fb_event_destroy(fb_event e);
fb_status fb_event_record(fb_event e, fb_stream s);
fb_status fb_event_elapsed_ms(fb_event start, fb_event stop, float *ms);
fb_status fb_event_synchronize(fb_event e);

const char *fb_status_string(fb_status s);
const char *fb_last_error(void);
uint32_t fb_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FUZZYBLUE_HIP_H_ */
