/*
 * ppm.h -- C ABI of the MI355X-native stochastic progressive photon mapping (SPPM) pass loop.
 *
 * The reference renders one SPPM pass per GLUT frame in m_display() (glslppm.cpp:214-536): eye
 * trace every MaxNumberOfBounces frames, max/min reduction of hit positions, photon trace, sum of
 * emitted flags, hash, scatter, correction, progressive update, draw.  Its GPU state lives in
 * RGBA32F textures bound to GL programs; there is no library API.  This header replaces that GL
 * texture ping-pong with three calls a render loop can make on one GPU:
 *
 *   ppm_create   -- replaces the texture/FBO set-up of main() (glslppm.cpp:692-795), the scene
 *                   upload of CGPURT::CreateTextures (gpurt.cpp:47-552) and RandomizeTextures
 *                   (glslppm.cpp:113-140).
 *   ppm_pass     -- replaces m_display() (glslppm.cpp:214-536), same stage order, asynchronous.
 *   ppm_readback -- replaces glReadPixels / reading the textures (glslppm.cpp:206, the draw of
 *                   re.fs at glslppm.cpp:506-513).
 *
 * The host scene layer (OBJ/MTL load of mesh.cpp, camera.cpp, the threaded multi-BVH of bvh.cpp)
 * is exposed as ppm_mesh_* and produces the flattened ppm_scene that ppm_create consumes.  A
 * caller that already holds the reference's TMesh/TBVH fills ppm_scene itself.
 *
 * Conventions: every int-returning call returns PPM_OK (0) or a negative PPM_E_* code and never
 * aborts; the message is available from ppm_last_error / ppm_mesh_last_error.  A context is used
 * by one thread at a time; distinct contexts may be driven from distinct threads.  Caller-owned
 * buffers are copied, never retained.  Pixel p = j*width + i with row j counted from the bottom,
 * the GL convention of gl_FragCoord (glslppm.cpp:120-124).
 */
#ifndef PPM_H
#define PPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPM_OK 0
#define PPM_E_INVAL (-1)   /* bad argument, size or state */
#define PPM_E_OOM (-2)     /* device or host allocation failed */
#define PPM_E_HIP (-3)     /* a HIP runtime call failed */
#define PPM_E_IO (-4)      /* file could not be read */
#define PPM_E_NODEV (-5)   /* no usable GPU */

/* BRDF codes of the material table (mesh.cpp:520-611, used by eyeraytrace.fs:448-539 and
 * photontrace.fs:567-669). */
#define PPM_BRDF_LIGHT (-1)
#define PPM_BRDF_DIFFUSE 0
#define PPM_BRDF_MIRROR 1
#define PPM_BRDF_GLASS 2
#define PPM_BRDF_PLASTIC 3
#define PPM_BRDF_GLOSSY_MIRROR 4
#define PPM_BRDF_GLOSSY_GLASS 5
#define PPM_BRDF_GLOSSY_PLASTIC 6

/* Look-at pinhole camera, TCamera::Set (camera.cpp:4-18) and the CameraParams uniform
 * (glslppm.cpp:270-273). */
typedef struct ppm_camera {
    float origin[3];
    float u[3], v[3], w[3];
    float half_width;   /* width * 0.5f  */
    float half_height;  /* height * 0.5f */
    float distance;     /* height / (2 tan(fov/2)) */
} ppm_camera;

/* One material, the two texels of TextureMaterials (gpurt.cpp:177-198): (rgb, eta), (brdf, g). */
typedef struct ppm_material {
    float color[3];
    float eta;
    int32_t brdf;
    float g;            /* 1 / ((2 / (1 - specularity))^2.71828 + 1), gpurt.cpp:194 */
} ppm_material;

/* Flattened scene: triangles, materials, light CDF and the threaded six-order BVH.
 * Node k is the k-th node of the canonical (face 0) pre-order of TBVH::Build (bvh.cpp:280-353).
 * node_links[(f * num_nodes + k) * 2 + {0,1}] are the hit and miss links of node k in the order
 * of cube face f (0:+x 1:-x 2:+y 3:-y 4:+z 5:-z), as canonical node indices, -1 for "done";
 * they are the texels of the six link tiles of textureBVHs (gpurt.cpp:408-473). */
typedef struct ppm_scene {
    int32_t num_triangles;
    const float* positions;       /* [T][3 vertices][xyz] */
    const float* normals;         /* [T][3 vertices][xyz] shading normals before packing */
    const int32_t* material;      /* [T] material index */
    int32_t num_materials;
    const ppm_material* materials;/* [M] */
    int32_t num_nodes;            /* 2T-1 */
    const float* node_bounds;     /* [N][6] min xyz, max xyz */
    const int32_t* node_triangle; /* [N] leaf triangle or -1 */
    const int32_t* node_links;    /* [6][N][2] */
    int32_t num_lights;
    const int32_t* light_triangle;/* [NL] triangles whose material is a light */
    const float* light_cdf;       /* [NL] normalised running area sum (mesh.cpp:456-488) */
    float light_area;             /* LightSummedArea */
    float bsphere[4];             /* SceneBSphere: bbox centre, half diagonal (glslppm.cpp:388-392) */
    ppm_camera camera;
    /* Textures (mesh.cpp:37-127, 656-669; gpurt.cpp:158-198).  texcoords holds each triangle
     * vertex's (u, v) for triangles of textured materials and 1e30 elsewhere; NULL when no
     * triangle is textured (then no lookup ever happens).  textures is the RGB8 volume
     * [num_materials][texture_size][texture_size][3]: a textured material's image resampled to
     * texture_size^2 by nearest texel, an untextured material's slice magenta (255, 0, 255). */
    const float* texcoords;       /* [T][3][2] or NULL */
    int32_t texture_size;         /* 512 (gpurt.cpp:170), 0 without textures */
    const uint8_t* textures;      /* [M][S][S][3] or NULL */
} ppm_scene;

/* Pass-loop parameters.  The reference couples Q = P = H = 512^2 (glslppm.cpp:36-40); here the
 * three sizes are independent. */
typedef struct ppm_params {
    int32_t image_width, image_height;   /* Q eye pixels / hit points */
    int32_t photon_width, photon_height; /* P photon slots (PhotonBufferSize^2 in the reference) */
    int32_t hash_width, hash_height;     /* H hash slots; hash_width a power of two, H <= 2^24 */
    int32_t max_path_length;             /* MaxNumberOfBounces: bounce cap and eye-pass cadence */
    float alpha;                         /* 0.7 (glslppm.cpp:488) */
    float initial_footprint;             /* 2.5 (glslppm.cpp:40) */
    float fixed_radius;                  /* > 0: initial radius used instead of the footprint rule */
    float focal_length;                  /* 13 (glslppm.cpp:31) */
    float aperture;                      /* 0 (glslppm.cpp:32) */
    uint32_t xorshift[4];                /* 123456789, 362436069, 521288629, 88675123 */
    int32_t device;                      /* HIP device ordinal */
    uint32_t flags;                      /* PPM_FLAG_* */
} ppm_params;

#define PPM_FLAG_NO_DRAW 1u     /* skip the per-pass RGBA8 draw (re.fs); readback still resolves */
#define PPM_FLAG_GRAPH 2u       /* ppm_pass replays each pass as a captured hipGraph (one for eye
                                   passes, one for photon-only passes) instead of 9 launches */

/* Buffers ppm_readback can copy out.  Element types and counts are given per entry; the byte
 * count passed must equal the buffer size exactly. */
typedef enum ppm_buf {
    PPM_BUF_RADIANCE = 0,       /* float[Q][4]: tau/(R^2 pi NumPhotons) + E, w = 1 (re.fs:35-38) */
    PPM_BUF_TONEMAPPED = 1,     /* float[Q][4]: 1-exp(-60000 L), sRGB (re.fs:41-47) */
    PPM_BUF_RGBA8 = 2,          /* uint8[Q][4]: the window image of the last pass */
    PPM_BUF_HIT_POSITION = 3,   /* float[Q][4]: QueryPositionTexture (pos, dir.x) */
    PPM_BUF_HIT_REFLECTANCE = 4,/* float[Q][4]: QueryReflectanceTexture (col, dir.y) */
    PPM_BUF_HIT_NORMAL = 5,     /* float[Q][4]: QueryNormalTexture (nrm, dir.z) */
    PPM_BUF_HIT_TRIANGLE = 6,   /* int32[Q][2]: triangle and material of the last eye hit, -1 miss */
    PPM_BUF_HIT_T = 7,          /* float[Q]: isect.w of the last eye trace */
    PPM_BUF_EYE_RNG = 8,        /* float[Q][4]: RandomEyeRayTexture */
    PPM_BUF_FLUX_RADIUS = 9,    /* float[Q][4]: QueryFluxRadiusTexture (tau, R) */
    PPM_BUF_EMISSION_COUNT = 10,/* float[Q][4]: QueryEmissionPhotonCountTexture (E, N) */
    PPM_BUF_PHOTON_POSITION = 11,  /* float[P][4]: (pos, nrm.x) */
    PPM_BUF_PHOTON_FLUX = 12,      /* float[P][4]: (flux, nrm.y); flux.x < 0 marks no deposit */
    PPM_BUF_PHOTON_DIRECTION = 13, /* float[P][4]: (dir, nrm.z) */
    PPM_BUF_PHOTON_RNG = 14,       /* float[P][4]: RandomPhotonTexture */
    PPM_BUF_PHOTON_STATE = 15,     /* int32[P][3]: triangle (-1 none), bounce count, emitted flag */
    PPM_BUF_PHOTON_SLOT = 16,      /* int32[P]: hash slot of each photon (PhotonIndexTexture.zw), -1 */
    PPM_BUF_SLOT_COUNT = 17,       /* uint32[H]: PhotonCorrectionTexture */
    PPM_BUF_SLOT_REP = 18,         /* int32[H]: representative photon (min index), -1 when none */
    PPM_BUF_COUNTERS = 19,         /* ppm_counters */
    PPM_BUF_GRID = 20              /* ppm_grid */
} ppm_buf;

typedef struct ppm_counters {
    uint64_t frame_count;      /* FrameCount: passes run */
    uint64_t eye_samples;      /* NumEyeSamples */
    uint64_t photons_emitted;  /* exact sum of emitted flags */
    uint64_t photon_segments;  /* P * passes */
    float num_photons_f;       /* NumPhotons as the reference accumulates it, in float */
    uint32_t reserved;
} ppm_counters;

typedef struct ppm_grid {
    float bbmin[3];            /* BBMin after expansion by the initial radius */
    float bbmax[3];
    float initial_radius;      /* InitialRadius */
    float grid_scale;          /* GridScale = 0.5 / InitialRadius */
    float hash_scale;          /* HashScale1 = 4194304 / GridScale */
} ppm_grid;

typedef struct ppm_ctx ppm_ctx;

/* Default parameters: the reference's constants at a given square size. */
void ppm_params_default(ppm_params* p, int32_t resolution);

/* Replaces the set-up of main() (glslppm.cpp:692-829): copies the scene to HBM, allocates all
 * pass state, seeds the per-pixel and per-photon generators from the xorshift state. */
int ppm_create(const ppm_scene* scene, const ppm_params* params, ppm_ctx** out);

/* Replaces m_display() (glslppm.cpp:214-536), run npasses times back to back on the context's
 * stream.  Returns once the work is queued.  Each pass's gather and draw (progressive.fs, re.fs)
 * run in the photon-stage launch of the pass after it, the last pass's at the end of the call, so
 * once the call's work is done every one of its passes is complete. */
int ppm_pass(ppm_ctx* ctx, uint32_t npasses);

/* ppm_pass for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  The same
 * passes with bit-identical results, launched stream by stream (never as graphs, whatever
 * PPM_FLAG_GRAPH says), with the passes' partial results (the eye kernel's per-block min/max of
 * hit positions, 24 bytes per tile of 256 pixels) in a static device array of 1 MiB per GPU
 * instead of the context's own buffer; the call allocates nothing and records no per-kernel
 * timing events.  Images above about 11M pixels (43,690 tiles) do not fit: PPM_E_INVAL.  Every
 * call on one GPU shares that array: calls whose passes overlap on the GPU (contexts on
 * different streams, or an eye stage a call started ahead that is still running; ppm_sync waits
 * for both) race and give wrong results. */
int ppm_pass_static(ppm_ctx* ctx, uint32_t npasses);

/* Waits for queued passes and copies one buffer to host memory. */
int ppm_readback(ppm_ctx* ctx, ppm_buf which, void* dst, size_t bytes);

/* Size in bytes of a readback buffer for this context, 0 for an unknown id. */
size_t ppm_buffer_bytes(const ppm_ctx* ctx, ppm_buf which);

/* Waits for queued passes. */
int ppm_sync(ppm_ctx* ctx);

/* Device time in milliseconds of the last ppm_pass call, measured with HIP events on the context
 * stream; -1 when not available. */
float ppm_last_pass_ms(ppm_ctx* ctx);

/* Per-kernel device timing with HIP events recorded around each launch on the context stream.
 * ppm_profile_enable(ctx, 1) clears the accumulators and starts recording every kind;
 * ppm_profile_enable(ctx, PPM_PROFILE_KINDS | 1 << kind | ...) records only those kinds (each event
 * costs the stream a few microseconds, so time only what is needed); 0 stops.  ppm_profile_read
 * waits for the streams and returns the accumulated milliseconds and launch counts per kernel kind.
 * No counterpart in the reference (it times whole frames with glutGet, glslppm.cpp:526-535). */
#define PPM_KERNEL_EYE 0      /* eyeraytrace.fs + max/min */
#define PPM_KERNEL_CLEAR 1    /* k_pass_begin: the pass's uniforms (the glClear of the hash and
                                 correction targets runs inside the photon stage of the pass
                                 before, on the next of three hash tables) */
#define PPM_KERNEL_PHOTON 2   /* photontrace.fs + sum + hash + scatter + correction, with the
                                 progressive.fs + re.fs of the pass before in the same launch */
#define PPM_KERNEL_GATHER 3   /* progressive.fs + re.fs of the last pass of a ppm_pass call */
#define PPM_KERNEL_KINDS 4
#define PPM_PROFILE_KINDS 0x100  /* ppm_profile_enable: the low bits select kernel kinds */
int ppm_profile_enable(ppm_ctx* ctx, int enable);
int ppm_profile_read(ppm_ctx* ctx, double* ms, uint64_t* launches);

/* Launch shapes.  No counterpart in the reference (GL rasterises one fragment per texel).  Every
 * shape gives bit-identical results; only the speed differs.  ppm_autotune runs real passes of
 * the context (they count as passes) under each candidate value of one knob at a time -- photon
 * walk block size, pixel tile shape, graph replay, walk stack depth, eye-stage overlap -- and keeps
 * the fastest (a new value must beat the kept one by 1%);
 * call it before the passes whose speed matters. */
typedef struct ppm_tuning {
    int32_t trace_block;       /* threads per block of the photon walk: 64, 128, ..., 512 */
    int32_t tile_wave_width;   /* a wave's eye/gather tile is w x 64/w pixels: 1, 2, 4, ..., 64 */
    uint32_t passes_per_graph; /* plain loop: 0 launches, 1 hipGraph per pass (PPM_FLAG_GRAPH);
                                  batch: passes per hipGraph, 0 for plain launches */
    int32_t walk_stack;        /* LDS stack entries per lane of the BVH walks, 0..32; deeper
                                  walks go on over the pre-order node records */
    int32_t eye_overlap;       /* 1: eye stages run on a second stream beside the photon stage
                                  (default in a batch, and in the plain loop for images of 2^18
                                  pixels or more; the plain loop starts each eye pass's stage
                                  right after the eye pass before it), 0: one stream */
    float ms_per_pass;         /* ppm_autotune / ppm_batch_autotune: time of the chosen shape */
} ppm_tuning;

int ppm_get_tuning(ppm_ctx* ctx, ppm_tuning* t);
int ppm_set_tuning(ppm_ctx* ctx, const ppm_tuning* t);
/* passes_per_trial 0: four times max_path_length. */
int ppm_autotune(ppm_ctx* ctx, uint32_t passes_per_trial, ppm_tuning* chosen);

void ppm_destroy(ppm_ctx* ctx);
const char* ppm_last_error(const ppm_ctx* ctx);

/* ---------------- batched pass loop ----------------
 *
 * No counterpart in the reference, which renders one scene per process (glslppm.cpp:664-832).  A
 * batch holds many jobs -- each one scene, camera, parameter set and SPPM state, exactly what
 * ppm_create holds -- resident in HBM, and ppm_batch_pass advances all of them with one launch per
 * stage of m_display (glslppm.cpp:214-536).  Each job keeps its own frame count, eye cadence,
 * XORShift stream and pass budget on the device; a job whose budget is spent is skipped by the
 * kernels.  Every job's buffers after n of its passes are bit-identical to those of the same
 * scene and parameters after ppm_pass(ctx, n).  With plain launches (passes_per_graph 0, the
 * default) each job's eye stage runs ahead of its eye pass on a second stream beside the photon
 * passes; otherwise the launch sequence is captured as a hipGraph of passes_per_graph passes and
 * replayed (captured again after ppm_batch_add/remove) and each eye stage runs in its eye pass.  Device
 * memory comes from a caching pool owned by the batch, so adding and removing jobs does not wait
 * for the passes in flight.  A batch is used by one thread at a time. */

typedef struct ppm_batch ppm_batch;

typedef struct ppm_batch_config {
    int32_t device;            /* HIP device ordinal; every job's params.device must match */
    uint32_t passes_per_graph; /* passes per captured hipGraph (1-64); 0: plain launches (default) */
} ppm_batch_config;

#define PPM_PASSES_UNLIMITED 0xffffffffu

void ppm_batch_config_default(ppm_batch_config* cfg);
int ppm_batch_create(const ppm_batch_config* cfg, ppm_batch** out);

/* Adds a job (copies the scene, allocates its state, seeds it like ppm_create) with a budget of
 * `passes` passes (PPM_PASSES_UNLIMITED: no limit).  Returns its id in *job. */
int ppm_batch_add(ppm_batch* b, const ppm_scene* scene, const ppm_params* params, uint32_t passes, int32_t* job);

/* Removes a job; its memory is reused once the passes queued so far have finished. */
int ppm_batch_remove(ppm_batch* b, int32_t job);

/* Sets a job's remaining pass budget, after the passes queued so far.  Synchronises. */
int ppm_batch_set_passes(ppm_batch* b, int32_t job, uint32_t passes);

/* Queues npasses batched passes; returns once they are queued. */
int ppm_batch_pass(ppm_batch* b, uint32_t npasses);

/* Waits for the queued passes. */
int ppm_batch_sync(ppm_batch* b);

/* Device time in milliseconds of the last ppm_batch_pass call (HIP events); -1 if unavailable. */
float ppm_batch_last_pass_ms(ppm_batch* b);

/* Passes a job has run and its remaining budget.  Synchronises. */
int ppm_batch_progress(ppm_batch* b, int32_t job, uint64_t* frames, uint32_t* passes_left);

/* ppm_readback and ppm_buffer_bytes of one job. */
int ppm_batch_readback(ppm_batch* b, int32_t job, ppm_buf which, void* dst, size_t bytes);
size_t ppm_batch_buffer_bytes(ppm_batch* b, int32_t job, ppm_buf which);

int32_t ppm_batch_num_jobs(const ppm_batch* b);

/* HBM held by the batch's pool (jobs in use plus cached blocks). */
uint64_t ppm_batch_device_bytes(const ppm_batch* b);

/* HBM a job of this scene and these parameters allocates in a batch (scene plus pass state), and
 * the pass state alone (what each further job on an already resident scene adds). */
uint64_t ppm_scene_footprint(const ppm_scene* scene, const ppm_params* params);
uint64_t ppm_state_footprint(const ppm_params* params);

/* Launch shapes of the whole batch (ppm_tuning); ppm_batch_autotune times real batched passes,
 * which count against the jobs' budgets. */
int ppm_batch_get_tuning(ppm_batch* b, ppm_tuning* t);
int ppm_batch_set_tuning(ppm_batch* b, const ppm_tuning* t);
int ppm_batch_autotune(ppm_batch* b, uint32_t passes_per_trial, ppm_tuning* chosen);

void ppm_batch_destroy(ppm_batch* b);
const char* ppm_batch_last_error(const ppm_batch* b);

/* ---------------- serving loop ----------------
 *
 * No counterpart in the reference.  A server owns one batch and a worker thread that runs it:
 * render jobs (a registered scene, a camera, parameters and a target pass count) arrive from
 * any thread at any time, wait until device memory and a slot are free, are admitted into the
 * running batch between steps of passes_per_step batched passes, keep their hit points, radii,
 * flux and generator state resident in HBM until their target is reached, and leave the batch
 * when their results have been copied out.  Admission and result copies overlap the passes of the
 * other jobs.  Every job's results equal those of the same scene, camera and parameters run
 * alone through ppm_create / ppm_pass for its target pass count, bit for bit.
 *
 * Scenes are registered once (packed on the calling thread) and shared by every job that names
 * them; a scene is uploaded with its first admitted job and stays resident while jobs use it. */

typedef struct ppm_server ppm_server;

typedef struct ppm_server_config {
    int32_t device;
    uint64_t hbm_budget;       /* device bytes for job state and scenes (default 200 GB) */
    int32_t max_jobs;          /* resident jobs at most (default 256) */
    uint32_t passes_per_step;  /* batched passes between admission points (default 8) */
    uint32_t passes_per_graph; /* the batch's hipGraph length (default 0: plain launches) */
} ppm_server_config;

#define PPM_JOB_QUEUED 0
#define PPM_JOB_RUNNING 1
#define PPM_JOB_DONE 2
#define PPM_JOB_FAILED 3

typedef struct ppm_job_info {
    int32_t state;             /* PPM_JOB_* */
    uint32_t target_passes;
    uint64_t passes_done;      /* as of the last step the server observed */
    uint64_t footprint;        /* device bytes of the job's state */
    double t_submit;           /* seconds on the server clock (ppm_server_now) */
    double t_admit;            /* -1 until admitted */
    double t_first_pass;       /* first observation of a finished first pass; -1 before */
    double t_done;             /* results copied out; -1 before */
} ppm_job_info;

typedef struct ppm_server_stats {
    int32_t queued, resident, done, failed;
    uint64_t device_bytes;     /* job state and resident scenes */
    uint64_t steps;            /* batched steps run */
    uint64_t photon_segments;  /* of finished jobs: P x target passes */
    double busy_seconds;       /* server clock time with at least one job resident */
} ppm_server_stats;

void ppm_server_config_default(ppm_server_config* cfg);
int ppm_server_create(const ppm_server_config* cfg, ppm_server** out);

/* Registers a scene (validated and packed on the calling thread); returns its id. */
int ppm_server_add_scene(ppm_server* s, const ppm_scene* scene, int64_t* scene_id);
/* Forgets a scene once no queued or running job uses it. */
int ppm_server_remove_scene(ppm_server* s, int64_t scene_id);

/* Queues a job.  camera may be NULL (the scene's own).  Returns at once with the job id. */
int ppm_server_submit(ppm_server* s, int64_t scene_id, const ppm_camera* camera, const ppm_params* params,
                      uint32_t target_passes, int64_t* job);

/* Waits until the job is done or failed, at most timeout_s seconds (< 0: no limit).  Returns
 * PPM_OK when done, 1 on time-out, the job's error code when it failed. */
int ppm_server_wait(ppm_server* s, int64_t job, double timeout_s);
int ppm_server_info(ppm_server* s, int64_t job, ppm_job_info* info);

/* A finished job's results: PPM_BUF_RADIANCE, PPM_BUF_RGBA8, PPM_BUF_FLUX_RADIUS,
 * PPM_BUF_EMISSION_COUNT or PPM_BUF_COUNTERS; bytes must equal the buffer size. */
int ppm_server_result(ppm_server* s, int64_t job, ppm_buf which, void* dst, size_t bytes);
/* Forgets a finished job and its host copies. */
int ppm_server_release(ppm_server* s, int64_t job);

int ppm_server_stats_get(ppm_server* s, ppm_server_stats* stats);
double ppm_server_now(const ppm_server* s);

/* Stops the worker after its current step; jobs not done fail with "server stopped". */
void ppm_server_destroy(ppm_server* s);
const char* ppm_server_last_error(const ppm_server* s);

/* ---------------- image files ----------------
 *
 * The reference only shows its image in a window (glslppm.cpp:506-526); these write the same
 * buffers to files.  Input rows are bottom-up, as every readback is.
 * ppm_write_png: PPM_BUF_RGBA8 as an 8-bit RGBA PNG (rows flipped to PNG's top-down order).
 * ppm_write_pfm: PPM_BUF_RADIANCE (float RGBA) as a colour PFM (RGB, little-endian, bottom-up). */
int ppm_write_png(const char* path, const uint8_t* rgba, int32_t width, int32_t height);
int ppm_write_pfm(const char* path, const float* rgba, int32_t width, int32_t height);

/* ---------------- host scene layer (mesh.cpp, camera.cpp, bvh.cpp) ---------------- */

typedef struct ppm_mesh ppm_mesh;

/* TMesh::LoadOBJ (mesh.cpp:491-693): triangles of an OBJ file with its MTL library, positions
 * scaled then offset, materials classified by the MTL rules of mesh.cpp:520-611. */
int ppm_mesh_load_obj(const char* path, const float position[3], float scale, ppm_mesh** out);

/* A mesh from arrays.  normals may be NULL (face normals are used, mesh.cpp:642-652). */
int ppm_mesh_from_arrays(int32_t num_triangles, const float* positions, const float* normals,
                         const int32_t* material, int32_t num_materials,
                         const ppm_material* materials, ppm_mesh** out);

/* Seeded synthetic scenes: kind 0 = Cornell room with a mirror box and a diffuse box (36
 * triangles); kind 1 = that room plus `count` jittered icospheres of subdivision `detail`
 * (diffuse, mirror, glass); kind 2 = value-noise heightfield of detail x detail cells with
 * `count` glass icospheres. */
#define PPM_SCENE_CORNELL 0
#define PPM_SCENE_SPHERES 1
#define PPM_SCENE_TERRAIN 2
int ppm_mesh_generate(int32_t kind, uint64_t seed, int32_t count, int32_t detail, ppm_mesh** out);

/* Texture coordinates of a mesh made from arrays, [T][3 vertices][u, v] (NULL clears them), and
 * the image of one material (RGB8, rows as stored in a PPM file, which makes the material
 * textured; NULL makes it untextured).  OBJ files get both from vt lines and map_Kd PPM images.
 * As in TMesh::LoadOBJ (mesh.cpp:656-669), a triangle uses its texture coordinates only when its
 * material is textured. */
int ppm_mesh_set_texcoords(ppm_mesh* mesh, const float* texcoords);
int ppm_mesh_set_texture(ppm_mesh* mesh, int32_t material, int32_t width, int32_t height, const uint8_t* rgb);

/* TCamera::Set (camera.cpp:4-18). */
int ppm_mesh_set_camera(ppm_mesh* mesh, const float origin[3], const float lookat[3],
                        int32_t width, int32_t height, float fov_degrees);

/* TBVH::Build (bvh.cpp:280-353) and the flattening of CGPURT::CreateTextures
 * (gpurt.cpp:269-552). */
int ppm_mesh_build(ppm_mesh* mesh);

/* The flattened scene; valid after ppm_mesh_build until the mesh is destroyed. */
const ppm_scene* ppm_mesh_scene(const ppm_mesh* mesh);

int32_t ppm_mesh_num_triangles(const ppm_mesh* mesh);
void ppm_mesh_destroy(ppm_mesh* mesh);
const char* ppm_mesh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* PPM_H */
