/*
 * pedalsky_hip.h -- C ABI of the MI355X (gfx950) sky + cloud transport library.
 *
 * This is the drop-in boundary for PedalSky's atmosphere + cloud passes. The
 * reference reaches those passes through the Rust `RenderSubsystem` trait
 * (crates/ps-core/src/subsystem.rs:66-118) implemented by
 * `AtmosphereSubsystem` (crates/ps-atmosphere/src/lib.rs:119-700) and
 * `CloudsSubsystem` (crates/ps-clouds/src/lib.rs:95-1279). Each function below
 * names the reference interface it replaces.
 *
 * Conventions
 *   - Plain C: POD structs, pointers and sizes, no exceptions cross the ABI.
 *   - Every call returns ps_status (0 = PS_OK). On failure ps_last_error()
 *     returns a thread-local message describing the last failing call.
 *   - A ps_ctx is not re-entrant: one context per thread, or serialise calls.
 *   - "device" pointers are HIP device allocations on the context's device;
 *     "host" pointers are ordinary CPU memory.
 *   - Units are metres, radians, and the reference's photometric proxy units.
 */
#ifndef PEDALSKY_HIP_H
#define PEDALSKY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PS_ABI_VERSION 1
/* MAX_CLOUD_LAYERS: crates/ps-clouds/src/params.rs:9, cloud_uniforms.wgsl:8 */
#define PS_MAX_CLOUD_LAYERS 8
#define PS_MAX_VIEWS 2

typedef int32_t ps_status;
enum {
    PS_OK = 0,
    PS_ERR_INVALID_ARGUMENT = 1,
    PS_ERR_DEVICE = 2,
    PS_ERR_OUT_OF_MEMORY = 3,
    PS_ERR_NOT_READY = 4,            /* a prerequisite upload or bake is missing */
    PS_ERR_OVERLAPPING_LAYERS = 5,   /* SynthesisError::OverlappingCloudLayers, ps-synthesis/src/state.rs:35-46 */
    PS_ERR_UNSUPPORTED = 6
};

/* ------------------------------------------------------------------------ */
/* Uniform structs, byte-identical to the reference's GPU uploads.           */
/* ------------------------------------------------------------------------ */

/* FrameUniformsGpu, crates/ps-core/src/frame_uniforms.rs:123-162 and
 * shaders/common/uniforms.wgsl:11-33. 448 bytes; matrices column-major. */
typedef struct ps_frame_uniforms {
    float view[16];
    float proj[16];
    float view_proj[16];
    float inv_view_proj[16];
    float camera_position_world[4];
    float camera_velocity_world[4];
    float sun_direction[4];          /* xyz unit direction, w angular radius (rad) */
    float sun_illuminance[4];        /* rgb illuminance proxy, w TOA lux */
    float lightning_illuminance[4];  /* rgb emission, w falloff radius (m) */
    float lightning_origin_world[4];
    float viewport_size[4];          /* w, h, 1/w, 1/h */
    float time_seconds;
    float simulated_seconds;
    uint32_t frame_index;
    float ev100;
    float prev_view_proj[16];
} ps_frame_uniforms;

/* AtmosphereParams (= WorldUniformsGpu), crates/ps-core/src/weather.rs:21-57,
 * shaders/common/uniforms.wgsl:39-58. 144 bytes. */
typedef struct ps_atmosphere_params {
    float planet_radius_m;
    float atmosphere_top_m;
    float rayleigh_scale_height_m;
    float mie_scale_height_m;
    float rayleigh_scattering[4];
    float mie_scattering[4];
    float mie_absorption[4];
    float mie_g;
    float _pad_after_mie_g[3];
    float ozone_absorption[4];
    float ozone_center_m;
    float ozone_thickness_m;
    float _pad_after_ozone[2];
    float ground_albedo[4];
    float haze_extinction_per_m[4];
} ps_atmosphere_params;

/* CloudParamsGpu, crates/ps-clouds/src/params.rs:31-140 and
 * shaders/clouds/cloud_uniforms.wgsl:23-83. 112 bytes. */
typedef struct ps_cloud_params {
    float sigma_s[3];
    float _pad_after_sigma_s;
    float sigma_a[3];
    float droplet_diameter_bias;
    uint32_t cone_light_sampling;
    float forward_bias;
    float detail_strength;
    float curl_strength;
    float powder_strength;
    float multi_scatter_a;
    float multi_scatter_b;
    float multi_scatter_c;
    float ambient_strength;
    float base_scale_m;
    float detail_scale_m;
    float weather_scale_m;
    uint32_t light_steps;
    uint32_t cloud_steps;
    uint32_t multi_scatter_octaves;
    uint32_t cloud_layer_count;
    uint32_t temporal_jitter;
    float wind_drift_strength;
    float wind_skew_strength;
    float diurnal_strength;
} ps_cloud_params;

/* CloudLayerGpu, crates/ps-core/src/weather.rs:98-134 and
 * shaders/clouds/cloud_uniforms.wgsl:85-106. 48 bytes. */
typedef struct ps_cloud_layer {
    float base_m;
    float top_m;
    float coverage;
    float density_scale;
    uint32_t cloud_type;             /* 0 Cu, 1 St, 2 Sc, 3 Ac, 4 As, 5 Ci, 6 Cs, 7 Cb */
    float shape_bias;
    float detail_bias;
    float anvil_bias;
    float droplet_diameter_um;
    float _pad_after_droplets[3];
} ps_cloud_layer;

/* LUT resolutions. The reference fixes them in
 * crates/ps-core/src/atmosphere_luts.rs:15-27 (256x64, 32x32, 192x108,
 * 32x32x64); they are a run-time choice here so the 32^3 froxel volume the
 * benchmark workloads use is expressible. */
typedef struct ps_lut_dims {
    int32_t transmittance_w, transmittance_h;
    int32_t multiscatter_w, multiscatter_h;
    int32_t skyview_w, skyview_h;
    int32_t ap_w, ap_h, ap_d;
} ps_lut_dims;

/* Noise storage formats (crates/ps-clouds/src/noise.rs:75-88 uses Rgba8Unorm /
 * Rg8Unorm; FP16 is the half-precision alternative). */
enum { PS_NOISE_UNORM8 = 0, PS_NOISE_FP16 = 1 };

/* Output pixel formats. */
enum { PS_PIXEL_F32 = 0, PS_PIXEL_F16 = 1 };

/* LUT selectors and bake flags (the four compute passes of
 * crates/ps-atmosphere/src/lib.rs:464-492). */
enum {
    PS_LUT_TRANSMITTANCE = 1,
    PS_LUT_MULTISCATTER = 2,
    PS_LUT_SKYVIEW = 4,
    PS_LUT_AERIAL_PERSPECTIVE = 8,
    PS_LUT_ALL = 15,
    /* multi_scattering tuning off: the MS LUT is cleared instead of baked
     * (crates/ps-atmosphere/src/lib.rs:527-556) */
    PS_LUT_MULTISCATTER_DISABLED = 16
};

/* Noise volume selectors for ps_read_noise. */
enum { PS_NOISE_BASE = 0, PS_NOISE_DETAIL = 1, PS_NOISE_CURL = 2 };

/* Render outputs; each pointer is a device buffer of W*H RGBA texels in
 * `format`, or NULL to skip that output.
 *   composite          sky pass followed by the dual-source cloud composite
 *                      (sky.wgsl:159-220, cloud_composite.wgsl:47-57):
 *                      rgb = cloud_lum + sky * cloud_trans, a = 1
 *   cloud_luminance    attachment 0 of cloud_march.wgsl fs_main (AP applied)
 *   cloud_transmittance attachment 1 of cloud_march.wgsl fs_main            */
typedef struct ps_image_out {
    void* composite;
    void* cloud_luminance;
    void* cloud_transmittance;
    int32_t format;
    int32_t _reserved;
} ps_image_out;

/* Algorithmic event counts of one render (for the roofline model). */
typedef struct ps_work_counters {
    uint64_t pixels;          /* pixels rendered */
    uint64_t layer_marches;   /* (pixel, layer) segments marched */
    uint64_t primary_steps;   /* primary march steps executed */
    uint64_t covered_steps;   /* steps past the coverage gate (base noise fetched) */
    uint64_t eroded_steps;    /* steps with cloud > 0 (curl + detail fetched) */
    uint64_t dense_samples;   /* steps with density > 1e-3 (light march + shading) */
    uint64_t light_steps;     /* light-march density evaluations */
    uint64_t light_covered;   /* light-march evaluations past the gates (base fetched) */
    uint64_t octaves;         /* multi-scatter octave evaluations */
    uint64_t ap_taps;         /* pixels whose cloud luminance received AP */
} ps_work_counters;

/* Per-kernel device times of the last timed frame sequence, averaged per
 * frame, in milliseconds (HIP events on the context stream). When a frame
 * bakes both the sky-view and the aerial-perspective LUTs they run in one
 * launch: skyview_ms then covers both and aerial_perspective_ms is ~0. */
typedef struct ps_frame_timing {
    float transmittance_ms;
    float multiscatter_ms;
    float skyview_ms;
    float aerial_perspective_ms;
    float render_ms;
    float frame_ms;           /* event-to-event time of one whole frame */
    int32_t frames;
    int32_t _reserved;
} ps_frame_timing;

typedef struct ps_ctx ps_ctx;

/* ------------------------------------------------------------------------ */
/* Context lifecycle. Replaces AtmosphereFactory::build / CloudsFactory::build
 * (crates/ps-atmosphere/src/lib.rs:684-699, crates/ps-clouds/src/lib.rs:1258-1278),
 * which allocate the subsystems' GPU resources.                             */
/* ------------------------------------------------------------------------ */
ps_status ps_ctx_create(int32_t device, ps_ctx** out_ctx);
void ps_ctx_destroy(ps_ctx* ctx);
/* Thread-local description of the last failure (never NULL). */
const char* ps_last_error(void);
int32_t ps_abi_version(void);
/* Use an existing hipStream_t (NULL = the context's own stream). */
ps_status ps_ctx_set_stream(ps_ctx* ctx, void* hip_stream);
ps_status ps_ctx_synchronize(ps_ctx* ctx);
/* LUT resolutions; reallocates the LUTs (atmosphere_luts.rs:15-27). */
ps_status ps_ctx_set_lut_dims(ps_ctx* ctx, const ps_lut_dims* dims);
ps_status ps_ctx_get_lut_dims(const ps_ctx* ctx, ps_lut_dims* out);

/* ------------------------------------------------------------------------ */
/* Per-frame state uploads. Replaces the bind-group-0/1 uniform writes
 * (FrameWorldBindings::write, crates/ps-core/src/bind_groups.rs:140) and
 * CloudsSubsystem::prepare (crates/ps-clouds/src/lib.rs:1105-1148).        */
/* ------------------------------------------------------------------------ */
ps_status ps_set_atmosphere(ps_ctx* ctx, const ps_atmosphere_params* atm);
ps_status ps_set_frame(ps_ctx* ctx, const ps_frame_uniforms* frame);
/* `layer_wind_xz`: for each layer, the (u east, v south) wind in m/s that
 * layer_wind_mps (cloud_density_common.wgsl:264-271) reads from the wind
 * field at the layer's mid altitude; NULL = calm. Layers must not overlap
 * (ps-synthesis/src/cloud_layers.rs check_non_overlap). */
ps_status ps_set_clouds(ps_ctx* ctx, const ps_cloud_params* params,
                        const ps_cloud_layer* layers, int32_t n_layers,
                        const float* layer_wind_xz);
/* CloudsSubsystem::set_march_steps (crates/ps-clouds/src/lib.rs:701-712):
 * the primary and light march step counts of the cloud parameters set by
 * ps_set_clouds, clamped to [8, 256] and [1, 16] as the reference clamps
 * them; they apply from the next render on (PS_ERR_NOT_READY before
 * ps_set_clouds). */
ps_status ps_set_march_steps(ps_ctx* ctx, uint32_t cloud_steps, uint32_t light_steps);

/* ------------------------------------------------------------------------ */
/* Sun position and top-of-atmosphere illuminance (host code). Replaces
 * ps_core::astro::sun_position / sun_position_precise
 * (crates/ps-core/src/astro/spa.rs:47-185) and the sun half of
 * WorldState::recompute (crates/ps-core/src/world.rs:95-111).             */
/* ------------------------------------------------------------------------ */
/* A UTC calendar instant (chrono DateTime<Utc> at whole seconds). */
typedef struct ps_utc {
    int32_t year, month, day, hour, minute, second;
} ps_utc;

/* SunPositionPrecise (spa.rs:29-45), plus the Julian day it was computed at. */
typedef struct ps_sun_position {
    double altitude_deg;             /* topocentric, above the horizon, no refraction */
    double azimuth_deg;              /* clockwise from north, [0, 360) */
    double zenith_deg;
    double distance_au;              /* Earth radius vector R */
    double right_ascension_deg;      /* geocentric, [0, 360) */
    double declination_deg;          /* geocentric */
    double apparent_sidereal_time_deg;
    double julian_day;
} ps_sun_position;

/* The WorldState sun fields (world.rs:30-37): SunPosition's FP32 values, the
 * world direction (+X east, +Y up, +Z south) and the TOA illuminance
 * 127500 lux / d^2. */
typedef struct ps_world_sun_state {
    float altitude_rad;
    float azimuth_rad;
    float distance_au;
    float direction_world[3];
    float toa_illuminance_lux;
} ps_world_sun_state;

/* julian_day_utc (spa.rs:194-216). */
double ps_julian_day(const ps_utc* utc);
/* delta_t_seconds (spa.rs:222-249): TT - UT for the year, seconds. */
double ps_delta_t_seconds(int32_t year);
/* sun_position_precise (spa.rs:60-185). Invalid dates or latitudes outside
 * [-90, 90] return PS_ERR_INVALID_ARGUMENT. */
ps_status ps_sun_position_precise(const ps_utc* utc, double latitude_deg, double longitude_deg,
                                  double elevation_m, ps_sun_position* out);
/* sun_position + WorldState::recompute's sun direction and TOA illuminance
 * (world.rs:95-100, altaz_to_world world.rs:106-111). */
ps_status ps_world_sun(const ps_utc* utc, double latitude_deg, double longitude_deg, double elevation_m,
                       ps_world_sun_state* out);
/* FrameUniforms::set_sun (frame_uniforms.rs:98-118) with the illuminance the
 * synthesised WeatherState carries, TOA lux on every channel
 * (ps-synthesis/src/state.rs:154). */
ps_status ps_frame_set_sun(ps_frame_uniforms* frame, const ps_world_sun_state* sun, float angular_radius_rad);

/* ------------------------------------------------------------------------ */
/* Cloud inputs. Replaces CloudNoise::bake (crates/ps-clouds/src/noise.rs:71-216)
 * and the weather textures ps_synthesis::synthesise uploads
 * (crates/ps-synthesis/src/state.rs:50-185).                               */
/* ------------------------------------------------------------------------ */
/* Bake the 128^3 base, 32^3 detail and 128^2 curl volumes on the GPU
 * (noise_base/detail/curl.comp.wgsl). Seed 0 reproduces the reference hash. */
ps_status ps_generate_noise(ps_ctx* ctx, uint32_t seed, int32_t format);
/* Upload pre-baked volumes: base 128^3 RGBA, detail 32^3 RGBA, curl 128^2 RG,
 * each texel in `format` (uint8 UNORM or IEEE binary16 bits). */
ps_status ps_upload_noise(ps_ctx* ctx, int32_t format, const void* base_rgba,
                          const void* detail_rgba, const void* curl_rg);
/* Read back a stored volume in its storage format (same layouts as upload). */
ps_status ps_read_noise(ps_ctx* ctx, int32_t which, void* host_out, size_t bytes);
/* 64x64 blue-noise ranks (R8Unorm, noise.rs:646-704). */
ps_status ps_upload_blue_noise(ps_ctx* ctx, const uint8_t* ranks_64x64);
/* The same rank map generated on the GPU into the context (replaces the
 * CPU generate_blue_noise(64) + texture upload of noise.rs:107-140,
 * 646-704): the reference's greedy void-and-cluster placement, bit-identical
 * to ps_generate_blue_noise. Asynchronous on the context's stream. */
ps_status ps_synthesise_blue_noise(ps_ctx* ctx);
/* Read the context's blue-noise ranks back (64x64 bytes; synchronises). */
ps_status ps_read_blue_noise(ps_ctx* ctx, uint8_t* ranks_64x64);
/* Weather map: w*h texels of `channels` binary16 values; channel 0 is the
 * coverage the march reads (cloud_march.wgsl:872-875). `cloud_type_grid`:
 * 128x128 u8 (255 = use the layer's type) or NULL for all-sentinel.
 * `overcast_mask`: 128x128 R8Unorm top-down mask (sky.wgsl:120-140) or NULL
 * for zero. */
ps_status ps_upload_weather(ps_ctx* ctx, const uint16_t* weather_f16, int32_t w,
                            int32_t h, int32_t channels,
                            const uint8_t* cloud_type_grid_128,
                            const uint8_t* overcast_mask_128);

/* Weather inputs synthesised on the GPU. Replaces WeatherMap::synthesise
 * (crates/ps-synthesis/src/weather_map.rs:46-101, with coverage_grid.rs:31-61
 * for the gridded R gate) and CloudTypeGrid::synthesise
 * (cloud_type_grid.rs:41-92). The grids are the data the scene's
 * coverage_grid files hold (row-major, grid_w x grid_h, spanning
 * [-extent/2, extent/2] on x and z). */
typedef struct ps_weather_synth {
    int32_t has_layers;              /* scene has cloud layers: R gate 1 (else 0) where there is no grid */
    float precip_mm_per_h;           /* scene.precipitation.intensity_mm_per_h */
    float ground_wetness;            /* surface.wetness.ground_wetness */
    const float* coverage_grid;      /* host f32 in [0, 1], or NULL */
    const uint8_t* type_grid;        /* host u8 cloud types (0..7; others mean "layer type"), or NULL */
    int32_t grid_w, grid_h;
    float grid_extent_m;
} ps_weather_synth;
/* Synthesise the 128x128 RGBA16F weather map (R coverage gate, G 0, B
 * cloud-base offset noise, A precipitation / wetness) and the cloud-type
 * grid, bit-identical to the reference's, and make them the context's
 * weather inputs (32 km tile; the march's weather_scale_m places it).
 * `overcast_mask_128` as in ps_upload_weather. */
ps_status ps_synthesise_weather(ps_ctx* ctx, const ps_weather_synth* synth, const uint8_t* overcast_mask_128);
/* Read back the last synthesised map (128*128*4 binary16) and, if
 * `cloud_types_128` is non-NULL, the cloud-type grid (128*128 u8). */
ps_status ps_read_weather(ps_ctx* ctx, uint16_t* rgba_f16, uint8_t* cloud_types_128);

/* ------------------------------------------------------------------------ */
/* Presentation: the frame the reference's swapchain (and its golden images)
 * holds, from the render's outputs and the depth buffer. Replaces, in frame
 * order: the ground pass (shaders/ground/pbr.wgsl, crates/ps-ground: GGX /
 * Smith / Lambert ground with Voronoi albedo, wet surface, puddles with
 * ripples, snow, cloud-shadowed sun and overcast ambient, then
 * L = lit * ap.a + ap.rgb at the surface distance, pbr.wgsl:653-666); the
 * cloud composite over it (cloud_composite.wgsl:47-57); godrays
 * (shaders/godrays/radial.wgsl + composite.wgsl, crates/ps-godrays: radial
 * blur toward the sun at half resolution, added to the HDR target); bloom
 * (shaders/bloom/{bright_pass,downsample,upsample}.wgsl, crates/ps-bloom: soft-knee bright pass, 3-level
 * 13-tap / 9-tap tent pyramid, added to the HDR target); and the tone-map
 * pass (shaders/postprocess/tonemap.wgsl:37-58 with the *Srgb swapchain
 * encode; crates/ps-postprocess/src/tonemap.rs).                           */
/* ------------------------------------------------------------------------ */
enum { PS_TONEMAP_ACES = 0, PS_TONEMAP_CLAMP = 1 };
/* The 8-bit output's target: the windowed app's *Srgb swapchain (the store
 * encodes the tone map's linear output to sRGB, tonemap.wgsl:57), or the
 * Rgba8Unorm target of the reference's headless harness, which stores the
 * linear output as it is (crates/ps-app/src/test_harness.rs:56-58, :72) —
 * the frames its golden images hold. */
enum { PS_TARGET_SRGB8 = 0, PS_TARGET_UNORM8 = 1 };
/* Ground shading: the caller's surface radiance (`ground_lit`, or none), or
 * the reference's ground pass evaluated on the depth buffer's surface. */
enum { PS_GROUND_RADIANCE = 0, PS_GROUND_PBR = 1 };
/* SurfaceMaterial::as_u32 (crates/ps-core/src/scene.rs:359-368, pbr.wgsl:135-139). */
enum { PS_MATERIAL_GRASS = 0, PS_MATERIAL_BARE_SOIL = 1, PS_MATERIAL_TARMAC = 2, PS_MATERIAL_SAND = 3,
       PS_MATERIAL_WATER_EDGE = 4 };

/* SurfaceParams, crates/ps-core/src/weather.rs:159-201 and
 * shaders/common/uniforms.wgsl:62-79. 64 bytes. */
typedef struct ps_surface_params {
    float visibility_m;
    float temperature_c;
    float dewpoint_c;
    float pressure_hpa;
    float wind_dir_deg;
    float wind_speed_mps;
    float ground_wetness;
    float puddle_coverage;
    float snow_depth_m;
    float puddle_start;
    float precip_intensity_mm_per_h;
    float precip_kind;
    float material;                  /* PS_MATERIAL_* as a float */
    float _pad[3];
} ps_surface_params;

typedef struct ps_present_params {
    float ev100;                     /* photographic EV at ISO 100 (tonemap.wgsl:50; also the frame's ev100 the
                                        bloom threshold is relative to, ps-bloom lib.rs:482-483) */
    int32_t tonemap;                 /* PS_TONEMAP_ACES (ACESFilmic) or PS_TONEMAP_CLAMP (Passthrough) */
    int32_t ground_model;            /* PS_GROUND_RADIANCE or PS_GROUND_PBR */
    /* GodraysTuning (crates/ps-core/src/config.rs:521-549) and the host's
     * runtime multiplier (ps-godrays lib.rs:337-339); godrays = 0 skips them */
    int32_t godrays;
    uint32_t godrays_samples;
    float godrays_decay;
    float godrays_intensity;         /* tuning intensity x intensity_scale */
    float godrays_bright_threshold;
    /* BloomTuning (config.rs:624-660); bloom = 0 skips it */
    int32_t bloom;
    float bloom_threshold_ev100;
    float bloom_intensity;
    float bloom_knee_ev;
    ps_surface_params surface;       /* read with PS_GROUND_PBR (WeatherState.surface) */
    int32_t target;                  /* PS_TARGET_SRGB8 (default) or PS_TARGET_UNORM8 */
} ps_present_params;
/* The reference's defaults: EV100 14, ACES, caller radiance, godrays on
 * (64 samples, decay 0.96, intensity 0.6, threshold 5000), bloom on
 * (threshold +2 EV, intensity 1, knee 0.5 EV), SurfaceParams::default(). */
void ps_default_present_params(ps_present_params* out);
/* `in`: the PS_PIXEL_F32 outputs of the ps_render call for this frame and
 * depth buffer. `ground_lit`: device W*H RGBA32F radiance of the opaque
 * surface before aerial perspective (the ground pass's `lit`), or NULL: the
 * composite is presented as it is; ignored with PS_GROUND_PBR, which shades
 * the surface itself (flat ground, normal +y, at the depth buffer's
 * positions). Outputs (device, either may be NULL but not both): W*H RGBA8
 * (alpha 255) encoded for `target`, and W*H RGBA32F HDR (the tone map's
 * input). Godrays and
 * bloom use scratch images the context keeps at the frame's size. */
ps_status ps_present(ps_ctx* ctx, const float* depth, const float* ground_lit, const ps_image_out* in,
                     const ps_present_params* params, uint8_t* out_srgb8, float* out_hdr);
/* Auto-exposure statistics of an HDR image (auto_exposure.comp.wgsl:31-73):
 * over the pixels on a 4-pixel grid, the sum of log2(max(luminance, 1e-6))
 * and the pixel count, written to `out_sum_count` (device, 2 floats). */
ps_status ps_exposure_stats(ps_ctx* ctx, const float* hdr_rgba32f, int32_t w, int32_t h, float* out_sum_count);
/* The EV100 that maps the mean log luminance to mid-grey
 * (crates/ps-postprocess/src/auto_exposure.rs:209-212). */
float ps_ev100_from_exposure_stats(float log_lum_sum, float pixel_count);

/* ------------------------------------------------------------------------ */
/* The four LUT passes (AtmosphereSubsystem::dispatch_pass,
 * crates/ps-atmosphere/src/lib.rs:494-584).                                */
/* ------------------------------------------------------------------------ */
ps_status ps_bake_luts(ps_ctx* ctx, uint32_t flags);
/* Copy one LUT to host as float RGBA (w*h[*d]*4 floats). */
ps_status ps_read_lut(ps_ctx* ctx, int32_t which, float* host_out, size_t floats);

/* ------------------------------------------------------------------------ */
/* Sky pass + cloud march + composite (atmosphere::sky, clouds::march,
 * clouds::composite: crates/ps-atmosphere/src/lib.rs:585-642,
 * crates/ps-clouds/src/lib.rs:736-994, 1054-1097). `depth` is a device
 * buffer of W*H reverse-Z NDC depths (0 = far, Depth32Float semantics of
 * crates/ps-core/src/framebuffer.rs) or NULL for all-far.                  */
/* ------------------------------------------------------------------------ */
ps_status ps_render(ps_ctx* ctx, const float* depth, const ps_image_out* out);
/* ps_render for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation); no
 * counterpart in the reference. Same output as ps_render bit for bit; it
 * stages the frame's parameter block in a fixed per-GPU scratch array (one
 * block, 4 KiB) instead of the context's allocated ring and calls no
 * allocating routine. Needs baked LUTs (ps_bake_luts) and the plain path:
 * half resolution, TAA and a perturbation volume return PS_ERR_UNSUPPORTED.
 * Every call on a GPU shares that GPU's scratch: calls that overlap on one
 * GPU (several contexts or streams at once) race; calls on one stream are
 * ordered. */
ps_status ps_render_noalloc(ps_ctx* ctx, const float* depth, const ps_image_out* out);
/* One whole frame: ps_bake_luts(flags) then ps_render. When `flags` rebakes
 * every LUT, the bake goes into a second LUT set on a side stream while the
 * previous frame still renders from the current one, and the sets swap
 * (ps_read_lut and later renders see this frame's LUTs): consecutive frames
 * overlap their LUT chain with the previous render. Results are identical
 * to the unpipelined sequence; under graph capture of the context's stream
 * the bake stays on that stream. */
ps_status ps_frame(ps_ctx* ctx, uint32_t lut_flags, const float* depth,
                   const ps_image_out* out);
/* Run `frames` frames back to back on the context stream and report average
 * per-kernel device times measured with HIP events. */
ps_status ps_time_frames(ps_ctx* ctx, int32_t frames, uint32_t lut_flags,
                         const float* depth, const ps_image_out* out,
                         ps_frame_timing* timing);
/* In-loop timing: between ps_timing_begin and ps_timing_end, every ps_frame
 * call (up to max_frames) records HIP events around its kernels; end
 * synchronises and reports the per-kernel averages. No reference
 * counterpart (the reference's GPU timestamp queries,
 * crates/ps-core/src/app.rs:195-227, serve the same purpose). */
ps_status ps_timing_begin(ps_ctx* ctx, int32_t max_frames);
ps_status ps_timing_end(ps_ctx* ctx, ps_frame_timing* timing);
/* Re-run the render with event counting on; fills the algorithmic counts. */
ps_status ps_count_work(ps_ctx* ctx, const float* depth, ps_work_counters* out);
/* ps_count_work that also records, per 8x8 wave tile (row-major), the
 * start and end of its work on the GPU's 100 MHz real-time clock into the
 * host array tile_times (2 entries per tile, capacity entries): the load
 * balance and tail of the render. No reference counterpart (diagnostics). */
ps_status ps_count_work_tiles(ps_ctx* ctx, const float* depth, ps_work_counters* out,
                              uint64_t* tile_times, int32_t capacity);
/* Half-resolution cloud march (`clouds.half_res_render`,
 * crates/ps-clouds/src/lib.rs:124-127 and 748-760; the march patch of
 * crates/ps-clouds/src/pipeline.rs:277-319; cloud_composite_halfres.wgsl):
 * with `enabled` != 0, ps_render and ps_frame march the clouds into a
 * (floor(W/2) x floor(H/2)) RT and composite at full resolution with a
 * Catmull-Rom upsample of the RT's luminance and transmittance (the
 * cloud_luminance / cloud_transmittance outputs are those upsampled
 * values). Off by default. ps_count_work, ps_render_train and batches run
 * at full resolution only (PS_ERR_UNSUPPORTED while it is on). */
ps_status ps_set_half_res(ps_ctx* ctx, int32_t enabled);
/* Cloud temporal anti-aliasing (`clouds.temporal_taa`, crates/ps-clouds/src/
 * lib.rs:128-142 and 773-829; shaders/clouds/cloud_taa.wgsl): with `enabled`
 * != 0, ps_render and ps_frame resolve each cloud RT (full or half
 * resolution) against the active view's history before the composite:
 * reprojection by the frame's inv_view_proj and prev_view_proj, 3x3
 * neighbourhood clamp, blend weight 1/8 (1 on the first frame of a history).
 * The caller passes `temporal_taa && !freeze_time` (the reference's gate).
 * An off -> on transition of a view, a change of the RT size and
 * ps_invalidate_taa_history discard history. Off by default. The cloud
 * outputs are the resolved values. ps_render_train and batches return
 * PS_ERR_UNSUPPORTED while it is on; ps_count_work counts the march only. */
ps_status ps_set_taa(ps_ctx* ctx, int32_t enabled);
/* `invalidate_taa_history` (crates/ps-clouds/src/lib.rs:726-734): every
 * view's next TAA frame starts from the current sample. */
ps_status ps_invalidate_taa_history(ps_ctx* ctx);
/* `set_active_view` (crates/ps-clouds/src/lib.rs:119-123, MAX_CLOUD_VIEWS
 * = 2 at lib.rs:69): the view (0 or 1) whose TAA history the next renders
 * read and write, for stereo hosts that render each eye in turn. */
ps_status ps_set_active_view(ps_ctx* ctx, int32_t view);

/* ------------------------------------------------------------------------ */
/* Many scenes per launch. No reference counterpart: the reference renders
 * one scene per App::frame (crates/ps-core/src/app.rs:294-322), running the
 * same passes per scene. A batch bakes every scene's own LUT set (one launch
 * per LUT kind for all scenes) and renders every wave tile of every scene
 * in one tile launch; the sequence can be captured once into a hipGraph and
 * replayed. Noise, weather map, cloud-type grid and blue noise are the
 * context's (shared by all scenes), as are the LUT sizes; the overcast mask
 * may be per scene. Scene parameters are read from the descriptors at create and
 * update time; depth and outputs are device pointers.                      */
/* ------------------------------------------------------------------------ */
typedef struct ps_scene_desc {
    const ps_frame_uniforms* frame;        /* viewport up to 4096 x 4096 */
    const ps_atmosphere_params* atmosphere;
    const ps_cloud_params* clouds;
    const ps_cloud_layer* layers;
    int32_t n_layers;
    int32_t _reserved;
    const float* layer_wind_xz;            /* 2 per layer, or NULL (calm) */
    const uint8_t* overcast_mask_128;      /* host 128x128 R8Unorm mask of this scene,
                                              or NULL for the context's */
    const float* depth;                    /* device W*H reverse-Z NDC, or NULL */
    ps_image_out out;                      /* device outputs (at least one) */
} ps_scene_desc;

/* Launch shape of a batch, chosen by ps_batch_autotune or by the caller. */
typedef struct ps_tuning {
    int32_t tile_w;          /* wave tile tile_w x (64 / tile_w) pixels: 4, 8 or 16 */
    int32_t persistent;      /* 1: one-wave workers pull tiles from an atomic counter;
                                0: one block per tile (hardware dispatch) */
    int32_t workers_per_cu;  /* persistent workers per compute unit (1..64) */
    int32_t order;           /* tile order: 0 scene order, 1 larger scenes first */
    int32_t grab;            /* persistent schedule step: consecutive tiles a worker takes per
                                counter increment (1..256; 0 means 1) */
} ps_tuning;

typedef struct ps_batch ps_batch;

ps_status ps_batch_create(ps_ctx* ctx, const ps_scene_desc* scenes, int32_t n, ps_batch** out);
/* New parameters for the same scenes (same count and resolutions); a
 * captured graph stays valid and reads them on its next replay. */
ps_status ps_batch_update(ps_batch* batch, const ps_scene_desc* scenes, int32_t n);
ps_status ps_batch_set_tuning(ps_batch* batch, const ps_tuning* tuning);
ps_status ps_batch_get_tuning(const ps_batch* batch, ps_tuning* tuning);
/* Time each candidate launch shape on this batch (`reps` runs each) and keep
 * the fastest; a captured graph is re-captured with it. */
ps_status ps_batch_autotune(ps_batch* batch, int32_t reps, ps_tuning* best);
/* Enqueue one batch run on the context stream (direct launches). */
ps_status ps_batch_run(ps_batch* batch);
/* Record the run sequence into a hipGraph; ps_batch_replay launches it. */
ps_status ps_batch_capture(ps_batch* batch);
ps_status ps_batch_replay(ps_batch* batch);
/* Run (or replay, when captured) `reps` times back to back; average device
 * time per run from HIP events on the context stream. */
ps_status ps_batch_time(ps_batch* batch, int32_t reps, float* ms_per_run);
void ps_batch_destroy(ps_batch* batch);
/* One-shot: create, run with `tuning` (NULL = default), wait, destroy. */
ps_status ps_render_batch(ps_ctx* ctx, const ps_scene_desc* scenes, int32_t n, const ps_tuning* tuning);

/* Step schedule of a batch: primary and light march steps per resolution
 * class (the batch's scenes grouped by viewport size, classes ordered by
 * pixel count, largest first). Each class's counts are what
 * CloudsSubsystem::set_march_steps (crates/ps-clouds/src/lib.rs:701-712)
 * would set for every scene of that size, clamped as it clamps them; 0 keeps
 * each scene's own counts. The reference has no batch: its host loop sets
 * one view's counts per frame to trade march detail for frame time. */
#define PS_MAX_STEP_CLASSES 8
typedef struct ps_step_schedule {
    int32_t n_classes;                           /* distinct viewport sizes (<= 8) */
    int32_t width[PS_MAX_STEP_CLASSES];
    int32_t height[PS_MAX_STEP_CLASSES];
    int32_t cloud_steps[PS_MAX_STEP_CLASSES];    /* [8, 256], or 0: the scenes' own */
    int32_t light_steps[PS_MAX_STEP_CLASSES];    /* [1, 16], or 0: the scenes' own */
    float batch_ms;                              /* autotune: device time of one run with it */
} ps_step_schedule;
ps_status ps_batch_get_step_schedule(const ps_batch* batch, ps_step_schedule* schedule);
/* Apply a schedule whose classes are the batch's (as get returns them);
 * counts are clamped. A captured graph stays valid (the counts travel with
 * the staged scene parameters). */
ps_status ps_batch_set_step_schedule(ps_batch* batch, const ps_step_schedule* schedule);
/* Choose the schedule for a device-time budget per batch run, as the
 * reference's adaptive-quality loop does for one view: starting from the
 * scenes' own counts (clamped), while a run (timed `reps` times with the
 * current launch shape) exceeds `budget_ms`, lower the primary steps of the
 * class with the most march work (pixels x steps over its scenes) by one
 * rung of 256 224 192 160 128 112 96 80 64 56 48 40 32 24 16 12 8; when every
 * class is at 8, lower the light steps likewise (16 12 10 8 6 5 4 3 2 1).
 * The chosen schedule stays applied and is written to `best` (with its
 * time); a budget below the all-minimum time ends there. */
ps_status ps_batch_autotune_steps(ps_batch* batch, float budget_ms, int32_t reps, ps_step_schedule* best);

/* ------------------------------------------------------------------------ */
/* Differentiation. No reference counterpart: the reference only renders
 * forward. These entry points differentiate the cloud march of ps_render
 * (cloud_march.wgsl:731-1054 with sample_density :432-530, march_to_light
 * :543-666, the shading :886-993 and the AP composite :1016-1048, plus the
 * composite of cloud_composite.wgsl:47-57).
 *
 * Two forward-identity extensions make the differentiated parameters
 * complete; their neutral values reproduce the reference exactly:
 *   - a per-layer extinction scale kappa_l multiplying the layer's sigma_t
 *     (Beer-Powder, the multi-scatter octaves and the step extinction);
 *   - an additive density perturbation volume: 128^3 FP32, x fastest, then
 *     the layer-normalised height h, then z; x and z span the weather tile
 *     (texel coordinate wx * 128 / weather_scale + 63.5, Repeat, with (wx,
 *     wz) the drift-offset position the weather map is read at), h is
 *     ClampToEdge (128 h - 0.5). Its trilinear value is added to the shaped
 *     cloud value before the final clamp, in the primary and the light-march
 *     density, past the altitude and coverage gates.                      */
/* ------------------------------------------------------------------------ */
/* `extinction_scale`: n host floats (layers >= n keep 1), NULL with n = 0
 * for all 1. `density_delta`: device 128^3 floats owned by the caller and
 * read by every later render (its contents may change between renders), or
 * NULL for none. */
ps_status ps_set_cloud_extensions(ps_ctx* ctx, const float* extinction_scale, int32_t n,
                                  const float* density_delta);
/* Transmittance checkpoint interval of ps_render_train, in primary steps
 * (1..32, default 16). */
ps_status ps_set_checkpoint_interval(ps_ctx* ctx, int32_t steps);
/* ps_render, and also store what ps_backward needs: the cloud transmittance
 * at the first step of every checkpoint chunk of every layer segment, and
 * per-pixel forward totals. The backward recomputes each chunk from its
 * checkpoint. */
ps_status ps_render_train(ps_ctx* ctx, const float* depth, const ps_image_out* out);
/* ps_bake_luts(lut_flags) then ps_render_train, with the LUT chain pipelined
 * as ps_frame pipelines it (a step of an optimisation loop bakes its LUTs
 * while the previous step's backward still runs). */
ps_status ps_frame_train(ps_ctx* ctx, uint32_t lut_flags, const float* depth, const ps_image_out* out);
/* Gradient outputs of ps_backward; both are ADDED to (the caller clears
 * them). */
typedef struct ps_cloud_grads {
    float* layers;          /* device, 24 floats: dL/d density_scale[8], dL/d extinction_scale[8],
                               dL/d coverage[8] (layer order of ps_set_clouds), or NULL */
    float* density_delta;   /* device 128^3 floats: dL/d density_delta, or NULL */
} ps_cloud_grads;
/* Backward of the last ps_render_train: the cotangents are device W*H RGBA
 * FP32 images (alpha ignored) of dL/d composite, dL/d cloud_luminance and
 * dL/d cloud_transmittance, each NULL for zero. The context's LUTs, inputs
 * and perturbation volume must be those of the training forward.
 * Perturbation-volume gradients need light_steps <= 32 (or the cone march).
 * Changing the extensions, the checkpoint interval, the LUT sizes, the noise
 * format or the weather-map size discards the training forward (then
 * PS_ERR_NOT_READY). */
ps_status ps_backward(ps_ctx* ctx, const float* g_composite, const float* g_luminance,
                      const float* g_transmittance, const ps_cloud_grads* grads);

/* ------------------------------------------------------------------------ */
/* Atmosphere calibration. No reference counterpart: the reference bakes the
 * LUT chain forward only (AtmosphereSubsystem::dispatch_pass,
 * crates/ps-atmosphere/src/lib.rs:494-584). The image differentiated here is
 * the sky + aerial-perspective image of the context's frame: the sky pass
 * (sky.wgsl:159-220, no cloud layers) where the depth buffer is far (0), and
 * where it holds a surface, that surface's radiance `ground_radiance`
 * attenuated and in-scattered by the AP froxels at its distance
 * (L = ground_radiance * ap.a + ap.rgb, the AP application of
 * pbr.wgsl:653-655 with the ground BRDF replaced by a constant). Frame, LUT
 * sizes and the non-differentiated atmosphere fields come from the context.
 * Coefficient sets are 14 floats: rayleigh_scattering[3], mie_scattering[3],
 * mie_absorption[3], ozone_absorption[3], rayleigh_scale_height_m,
 * mie_scale_height_m.                                                      */
/* ------------------------------------------------------------------------ */
/* The image for the context's atmosphere: W*H RGBA FP32 into device `out`.
 * Bakes its own LUT set (the context's LUTs are not touched). */
ps_status ps_sky_ap_image(ps_ctx* ctx, const float* depth, float ground_radiance, float* out);
/* For n coefficient sets (device n*14 floats): bake each set's LUT chain and
 * its tangents, render the image, and write to device `out` (n*16 floats)
 * per set the loss sum over pixels p and RGB c of w_p (image - target)^2
 * (out[16 s + 14]) and its gradient with respect to the 14 coefficients
 * (out[16 s + 0..13]). `target`: device W*H RGBA FP32 (alpha ignored);
 * `weight`: device W*H FP32 per-pixel weights w_p, or NULL for 1. `image`
 * (device W*H RGBA, or NULL) receives set 0's image. */
ps_status ps_atmosphere_grad(ps_ctx* ctx, int32_t n, const float* coeffs, const float* depth,
                             float ground_radiance, const float* target, const float* weight, float* out,
                             float* image);

/* ------------------------------------------------------------------------ */
/* Host-side synthesis helpers (CPU, no device work).                       */
/* ------------------------------------------------------------------------ */
/* AtmosphereParams::default (weather.rs:60-84), CloudParamsGpu::default
 * (params.rs:142-213). */
void ps_default_atmosphere(ps_atmosphere_params* out);
void ps_default_cloud_params(ps_cloud_params* out);
/* remap_coverage_to_visible_band, ps-synthesis/src/cloud_layers.rs:98-120 */
float ps_remap_coverage(float scene_coverage);
/* default_density_scale / default_droplet_diameter_um, ps-core/src/scene.rs:293-354 */
float ps_default_density_scale(uint32_t cloud_type);
float ps_default_droplet_diameter_um(uint32_t cloud_type);
/* synthesise_cloud_layers for one scene layer, cloud_layers.rs:44-73. A NaN
 * density_scale / anvil_bias / droplet_diameter_um selects the per-type
 * default. */
void ps_synthesise_cloud_layer(uint32_t cloud_type, float base_m, float top_m,
                               float scene_coverage, float density_scale,
                               float shape_bias, float detail_bias,
                               float anvil_bias, float droplet_diameter_um,
                               ps_cloud_layer* out);
/* check_non_overlap, cloud_layers.rs:134-148: returns 0 when disjoint, else
 * 1 and the first overlapping pair in out_i/out_j. */
int32_t ps_check_layers_overlap(const ps_cloud_layer* layers, int32_t n,
                                int32_t* out_i, int32_t* out_j);
/* The wind-field tap layer_wind_mps performs for a layer spanning
 * [base_m, top_m] under the procedural profile of
 * ps-synthesis/src/wind_field.rs:71-178 (power law + Ekman veer, f16
 * storage, linear filtering with repeat addressing). */
void ps_layer_wind(float surface_speed_mps, float wind_dir_deg, float base_m,
                   float top_m, float out_xz[2]);
/* Uniform top-down overcast mask value, density_mask.rs:134-165 with the
 * CPU NDF of ps-synthesis/src/ndf.rs:20-58. */
uint8_t ps_overcast_mask_value(const ps_cloud_layer* layers, int32_t n);
/* TopDownMask::synthesise (crates/ps-synthesis/src/density_mask.rs:47-83):
 * the 128x128 overcast mask over 32 km for a scene with a coverage grid
 * (grid_h rows of grid_w host floats over grid_extent_m, the scene's
 * [clouds.coverage_grid]); grid NULL gives the uniform mask of
 * ps_overcast_mask_value. */
void ps_overcast_mask(const ps_cloud_layer* layers, int32_t n, const float* grid, int32_t grid_w, int32_t grid_h,
                      float grid_extent_m, uint8_t* out_128x128);
/* generate_blue_noise(64), crates/ps-clouds/src/noise.rs:646-704. */
void ps_generate_blue_noise(uint8_t* out_64x64);

#ifdef __cplusplus
}
#endif

#endif /* PEDALSKY_HIP_H */
