/*
 * pf_hip.h — C ABI of the MI355X (gfx950) power-diagram ray tracer.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t
 * (passed as void*), enqueues its kernels on that stream and returns an int
 * status (PF_OK = 0).  Nothing here allocates device memory except where a
 * comment says so; the caller owns every input and output buffer.
 *
 * Reference interface replaced by each entry point is cited as file:line of
 * IsaacSuo/dynamic-powerfoam.  The reference exposes no C ABI: its boundary
 * is a set of Python classes that launch Warp kernels on torch's current
 * stream (powerfoam/warp_interop.py:5-24).  The entry points below are the
 * kernel launches those classes make, one for one.
 */
#ifndef PF_HIP_H
#define PF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PF_OK                 0
#define PF_ERR_ARG            1   /* bad pointer, size or enum              */
#define PF_ERR_DTYPE          2   /* unsupported attribute / face dtype     */
#define PF_ERR_HIP            3   /* a HIP runtime call failed              */
#define PF_ERR_OVERFLOW       4   /* a caller-sized buffer is too small     */

/* per-ray / per-call status bits written into a caller-provided int       */
#define PF_STATUS_STEP_CAP    1   /* a walk hit max_steps (reference: none,
                                     raytrace.py:163 loops forever)         */
#define PF_STATUS_LIST_CAP    2   /* a tile list exceeded its buffer         */
#define PF_STATUS_CYCLE       4   /* a walk returned to a (cell, log T) state
                                     it held before, so it would repeat
                                     forever (raytrace.py:163 never ends on
                                     such a ray); the ray stops there      */
#define PF_STATUS_BAD_RECORD  8   /* a walk record named a cell outside the
                                     scene, or its stamp names another
                                     scene, size or set of cameras (a
                                     record of another walk); the segment,
                                     or the whole record, was skipped       */

/* ---- dtypes ------------------------------------------------------------ */
#define PF_F32 0
#define PF_F16 1
#define PF_F16_PACKED 2   /* walk faces only: (E,3) 32-bit words per edge,
                             [half dp.x | half dp.y], [half dp.z | half dpm],
                             [int32 neighbour]; one 12-byte record per face,
                             so a face and its neighbour arrive in one load
                             (the values are those of PF_F16)             */

/* Pinhole camera, the fields of WarpCamera used on the pinhole path
 * (powerfoam/camera.py:8-17).  start_cell is the walk's first cell
 * (benchmark.py:320-324); rasterizer entry points ignore it.             */
typedef struct pf_camera {
    float eye[3];
    float right[3];
    float up[3];
    int32_t start_cell;
    int32_t width;
    int32_t height;
} pf_camera;                       /* 48 bytes */

/* A resident scene: device pointers in the reference's storage layout
 * (raytrace.py:274-280 packs spheres = [points, radii] and
 * nsigmas = [normals, density]; texel arrays are (N,K,3)/(N,K)).          */
typedef struct pf_scene {
    const float   *spheres;           /* (N,4) f32                          */
    const void    *nsigmas;           /* (N,4) attr dtype                   */
    const void    *texel_sites;       /* (N,K,3) attr dtype                 */
    const void    *texel_rgb;         /* (N,K,3) attr dtype                 */
    const void    *texel_height;      /* (N,K) attr dtype                   */
    const int32_t *adjacency;         /* (E,)   CSR neighbour indices       */
    const int32_t *adjacency_offsets; /* (N+1,) CSR row offsets             */
    const void    *faces;             /* (E,4) face dtype: walk faces
                                         (dp, dpm) or raster faces (dp, r_j) */
    int64_t num_cells;
    int64_t num_edges;
    int32_t num_texel_sites;          /* K, 1..16                           */
    int32_t attr_dtype;               /* PF_F32 | PF_F16                    */
    int32_t face_dtype;               /* PF_F32 | PF_F16 | PF_F16_PACKED    */
    int32_t reserved;
    int64_t texel_rgb_cam_stride;     /* elements between cameras' texel_rgb
                                         (view-dependent colour); 0: shared */
} pf_scene;

/* Record of each ray's contributing segments (segments with positive
 * length after the face and texel bounds), written by pf_walk_forward for
 * pf_walk_backward_recorded / pf_walk_backward_grouped to read instead of
 * replaying the walk.  No reference counterpart: the reference has no walk
 * backward.
 * Slots: one per lane of the walk's 8x8 tiles, slot = tile * 64 + lane,
 * tiles numbered camera-major then column-major over ceil(H/8) x ceil(W/8);
 * pf_walk_record_slots gives their number S.
 * segments: cap * S 16-byte records {cell, t_near bits, t_far bits,
 *           near face id | far face id << 16}; segment k of slot s is
 *           record ((k / 8) * S + s) * 8 + k % 8, so eight consecutive
 *           segments of a ray fill one 128-byte line.
 * counts:   (S) int32, each ray's number of contributing segments; rays
 *           with more than cap keep only their first cap, and the backward
 *           replays those rays instead.
 * cap:      a positive multiple of 8.
 * partials: NULL, or cap * S float4 {T before the segment, the ray's
 *           colour after it}, indexed as segments; required by
 *           pf_walk_backward_grouped.
 * header:   NULL, or 8 int32 that pf_walk_forward stamps with the scene's
 *           cell and edge counts, the cap, the slot count and a hash of the
 *           camera words; a backward given a record whose stamp differs
 *           from its own arguments reports PF_STATUS_BAD_RECORD and adds no
 *           gradient.
 * 16 * cap * S bytes of device memory for segments (2.1 GB at cap 64,
 * 1920x1080), as much again for partials.                                  */
typedef struct pf_walk_record {
    void    *segments;
    int32_t *counts;
    int32_t  cap;
    int32_t  reserved;
    void    *partials;
    int32_t *header;
} pf_walk_record;

/* Per-ray outputs of the walk.  Any pointer may be NULL except color.
 * Layout (B,H,W[,3]) row-major, B = number of cameras.                    */
typedef struct pf_walk_out {
    float    *color;      /* (B,H,W,3) composited RGB (raytrace.py:241)     */
    float    *log_t;      /* (B,H,W)   final log transmittance              */
    float    *depth;      /* (B,H,W)   depth where T crosses depth_quantile
                             (rasterize.py:871-877 semantics)               */
    int32_t  *visits;     /* (B,H,W)   number of cells visited               */
    uint64_t *seq_hash;   /* (B,H,W)   FNV-1a-64 of the visited cell ids     */
    int32_t  *seq;        /* (B,H,W,seq_cap) first seq_cap visited cells     */
    int32_t   seq_cap;
    int32_t   reserved;
    int32_t  *status;     /* one int, PF_STATUS_* bits OR-ed in (may be NULL)*/
    uint64_t *alg_bytes;  /* (B,H,W) algorithmic bytes the ray needed: per
                             visit sphere+CSR bounds+[n,sigma] and per face
                             index+face; per shaded segment K texels     */
    const pf_walk_record *record;  /* NULL, or the record to write        */
    int32_t  *ray_status; /* (B,H,W)   each ray's PF_STATUS_* bits (may be
                             NULL), e.g. to count the rays the loop guard
                             stopped                                        */
} pf_walk_out;

/* Gradients produced by the walk backward; all f32, accumulated (+=) so the
 * caller zeroes them.  Shapes follow RasterGradFn's outputs
 * (rasterize.py:1911-1915).                                                */
typedef struct pf_grads {
    float *spheres;       /* (N,4) d/d[points, radii]                       */
    float *nsigmas;       /* (N,4) d/d[normals, density]                    */
    float *texel_sites;   /* (N,K,3)                                        */
    float *texel_rgb;     /* (N,K,3)                                        */
    float *texel_height;  /* (N,K)                                          */
    int64_t texel_rgb_cam_stride; /* elements between cameras' texel_rgb
                                     gradients; 0: one shared array        */
} pf_grads;

/* Library / device information.  Returns the HIP device count in *n.      */
int pf_device_count(int32_t *n);
/* Human-readable message for a status code (static storage).             */
const char *pf_status_string(int32_t code);

/* ---- walk (RayTracer) --------------------------------------------------- */

/* Power-cell walk, forward.  Replaces RayTracer.benchmark's launch of
 * benchmark_kernel (powerfoam/raytrace.py:115-241, launch :289-307) and
 * extends it to a batch of cameras and the extra per-ray outputs above.
 * cams: device array of num_cams pf_camera, all of width x height.        */
int pf_walk_forward(const pf_scene *scene, const pf_camera *cams,
                    int32_t num_cams, int32_t width, int32_t height,
                    float transmittance_threshold, float depth_quantile,
                    int32_t max_steps, const pf_walk_out *out, void *stream);

/* flags of pf_walk_backward */
#define PF_BWD_EXACT_TEXEL    1   /* keep every texel and the radius term of
                                     the texel weights (default: the
                                     reference's approximations,
                                     texture.py:118-119, :158-159)          */

/* Power-cell walk, backward by checkpoint and replay.  No reference
 * counterpart: RayTracer defines only a forward benchmark_kernel
 * (raytrace.py:115-241) and no backward.  The
 * forward's per-ray color and log_t are the checkpoint; each ray's cell
 * sequence is replayed from its start cell and the gradients of
 * L(color, 1-exp(log_t)) are accumulated into grads.  With K = 8 each
 * 8x8 tile's wavefront keeps per-cell gradient rows in an LDS cache and
 * adds a row to grads once per residency (one atomic per element); other
 * K use per-segment atomics.  Summation order is not fixed, so gradients
 * agree with a fixed-order sum to rounding, not bit for bit.
 * grad_color (B,H,W,3), grad_log_t (B,H,W) (either may be NULL).
 * alg_bytes (B,H,W) or NULL: per-ray algorithmic bytes (the replay's reads
 * plus the gradient read-modify-writes).                                   */
int pf_walk_backward(const pf_scene *scene, const pf_camera *cams,
                     int32_t num_cams, int32_t width, int32_t height,
                     float transmittance_threshold, int32_t max_steps,
                     const float *color, const float *log_t,
                     const float *grad_color, const float *grad_log_t,
                     const pf_grads *grads, int32_t flags, int32_t *status,
                     uint64_t *alg_bytes, void *stream);

/* Number of record slots of a walk over num_cams cameras of width x
 * height: num_cams * ceil(height/8) * ceil(width/8) * 64.                  */
int64_t pf_walk_record_slots(int32_t num_cams, int32_t width, int32_t height);

/* Power-cell walk, backward from the forward's record (pf_walk_out.record
 * of the pf_walk_forward call with the same scene, cameras and
 * threshold).  No reference counterpart, as pf_walk_backward.  Rays the
 * record holds whole are differentiated from their recorded segments with
 * pf_walk_backward's per-segment arithmetic and gradient cache, without
 * reading a face row; the others are replayed as pf_walk_backward does.
 * The results equal pf_walk_backward's up to the order of float additions.
 * Other arguments as pf_walk_backward.                                    */
int pf_walk_backward_recorded(const pf_scene *scene, const pf_camera *cams,
                              int32_t num_cams, int32_t width, int32_t height,
                              float transmittance_threshold, int32_t max_steps,
                              const float *color, const float *log_t,
                              const float *grad_color, const float *grad_log_t,
                              const pf_walk_record *record,
                              const pf_grads *grads, int32_t flags,
                              int32_t *status, uint64_t *alg_bytes,
                              void *stream);

/* Bytes of device workspace pf_walk_backward_grouped needs for a record
 * of cap segments per ray over num_cams cameras of width x height and a
 * scene of num_cells cells; 0 if the sizes are out of range.              */
int64_t pf_walk_grouped_workspace_bytes(int64_t num_cells, int32_t num_cams,
                                        int32_t width, int32_t height,
                                        int32_t cap);

/* Power-cell walk, backward from the forward's record, grouped by cell.
 * No reference counterpart, as pf_walk_backward; it replaces the atomic
 * accumulation of the reference's raster backward (rasterize.py:1507-1516,
 * texture.py:125-171) for the walk.  The record must carry partials (the
 * transmittance before and the colour after each segment), which make every
 * segment's gradients independent of the rest of its ray.  The recorded
 * segments are counted and sorted by cell on the device (count, scan,
 * scatter), then differentiated in cell order with every lane busy; the
 * terms of equal cells are summed across lanes and added to grads once per
 * run of a cell.  Rays the record could not hold whole are replayed as
 * pf_walk_backward does, in the same launch.  K = 8 texels (other K: use
 * pf_walk_backward_recorded).  Results equal pf_walk_backward_recorded's up
 * to the order of float additions.
 * workspace: device memory of pf_walk_grouped_workspace_bytes bytes
 * (contents need not be initialised).  Other arguments as
 * pf_walk_backward_recorded; alg_bytes (B,H,W) per ray, or NULL.           */
int pf_walk_backward_grouped(const pf_scene *scene, const pf_camera *cams,
                             int32_t num_cams, int32_t width, int32_t height,
                             float transmittance_threshold, int32_t max_steps,
                             const float *color, const float *log_t,
                             const float *grad_color, const float *grad_log_t,
                             const pf_walk_record *record,
                             const pf_grads *grads, int32_t flags,
                             void *workspace, int64_t workspace_bytes,
                             int32_t *status, uint64_t *alg_bytes,
                             void *stream);

/* pf_walk_backward_grouped without a workspace argument: the library's
 * static device scratch array (pf_walk_grouped_scratch_bytes bytes, 64 MiB,
 * one per GPU, held until the process ends) serves as the workspace, and
 * the call allocates nothing.  No counterpart in the reference; it is for
 * callers that cannot allocate (code under HIP graph capture, serving
 * loops, harnesses that forbid allocation).  Sizes whose workspace
 * (pf_walk_grouped_workspace_bytes) exceeds the scratch array return
 * PF_ERR_OVERFLOW before any device work.  All calls on one GPU share its
 * array: calls that overlap on one GPU race.  Other arguments and results
 * as pf_walk_backward_grouped.                                             */
int pf_walk_backward_grouped_noalloc(const pf_scene *scene,
                                     const pf_camera *cams, int32_t num_cams,
                                     int32_t width, int32_t height,
                                     float transmittance_threshold,
                                     int32_t max_steps, const float *color,
                                     const float *log_t,
                                     const float *grad_color,
                                     const float *grad_log_t,
                                     const pf_walk_record *record,
                                     const pf_grads *grads, int32_t flags,
                                     int32_t *status, uint64_t *alg_bytes,
                                     void *stream);

/* Size in bytes of the static scratch array of
 * pf_walk_backward_grouped_noalloc.                                        */
int64_t pf_walk_grouped_scratch_bytes(void);

/* Checkpoint-and-replay backward through the cell-grouped pipeline.  No
 * reference counterpart, as pf_walk_backward: RayTracer defines only a
 * forward (raytrace.py:115-241).  No record from the forward: color /
 * log_t are the checkpoint, and the walk is replayed
 * band_tiles tiles (8x8 rays) at a time, in the cameras' tile order, by
 * the recording forward (cap record_cap, a multiple of 8) into a
 * band-sized record that the grouped pipeline sorts by cell and
 * differentiates.  Gradients are added into grads (zero them first).
 * workspace: device memory of at least
 * pf_walk_replay_grouped_workspace_bytes(...) bytes for the same sizes;
 * it holds one band, so its size follows band_tiles.  K = 8 texels.
 * Gradients equal pf_walk_backward's up to the order of float additions. */
int pf_walk_backward_replay_grouped(const pf_scene *scene,
                                    const pf_camera *cams, int32_t num_cams,
                                    int32_t width, int32_t height,
                                    float transmittance_threshold,
                                    int32_t max_steps, const float *color,
                                    const float *log_t,
                                    const float *grad_color,
                                    const float *grad_log_t,
                                    const pf_grads *grads, int32_t flags,
                                    int32_t record_cap, int32_t band_tiles,
                                    void *workspace, int64_t workspace_bytes,
                                    int32_t *status, void *stream);

/* Workspace bytes of pf_walk_backward_replay_grouped (0: sizes out of
 * range).                                                                  */
int64_t pf_walk_replay_grouped_workspace_bytes(int64_t num_cells,
                                               int32_t num_cams,
                                               int32_t width, int32_t height,
                                               int32_t cap,
                                               int32_t band_tiles);

/* Walk face precompute (benchmark.py:295-302): for every CSR edge i->j,
 * faces[e] = half(p_j - p_i, pm_j - pm_i) with pm = 0.5(|p|^2 - r^2).
 * face_dtype PF_F16 (reference), PF_F32 or PF_F16_PACKED.                  */
int pf_walk_faces(const float *points, const float *radii, int64_t num_cells,
                  const int32_t *adjacency, const int32_t *adjacency_offsets,
                  int64_t num_edges, void *faces, int32_t face_dtype,
                  void *stream);

/* Repacks reference-layout walk faces ((E,4) f16, the adjacency_diff that
 * benchmark.py:295-302 builds) and the adjacency into PF_F16_PACKED
 * records (E,3) int32.  No reference counterpart (a layout change only).  */
int pf_walk_pack_faces(const void *faces_f16, const int32_t *adjacency,
                       int64_t num_edges, void *packed, void *stream);

/* Start cell (benchmark.py:320-324): for each camera,
 * argmin_i |p_i - eye|^2 - r_i^2 (first index on ties), written into
 * cams[b].start_cell on the device.  workspace: caller-provided device
 * buffer of at least 8*num_cams bytes (contents need not be initialised).  */
int pf_start_cells(const float *points, const float *radii, int64_t num_cells,
                   pf_camera *cams, int32_t num_cams, void *workspace,
                   void *stream);

/* ---- tile rasterizer (Rasterizer) ----------------------------------------- */

/* Outputs of Rasterizer._forward (rasterize.py:1639-1679, 1850-1860);
 * (H,W[,3|Q]) row-major, contrib/point_err (N,) accumulated (+=).         */
typedef struct pf_raster_out {
    float *color;            /* (H,W,3)                                     */
    float *log_t;            /* (H,W)                                       */
    float *normal_distance;  /* (H,W)                                       */
    float *normal;           /* (H,W,3)                                     */
    float *quantile_depths;  /* (H,W,Q), caller fills with -1; may be NULL  */
    float *err;              /* (H,W) surface objective only; may be NULL   */
    float *contrib;          /* (N,)                                        */
    float *point_err;        /* (N,) when return_point_err; may be NULL     */
} pf_raster_out;

/* Incoming gradients of RasterGradFn.backward (rasterize.py:1865-1877),
 * after grad_log_t = -grad_opacity * exp(log_t) (:1909).  color required. */
typedef struct pf_raster_grad_in {
    const float *color;            /* (H,W,3)                               */
    const float *log_t;            /* (H,W) or NULL                         */
    const float *normal_distance;  /* (H,W) or NULL                         */
    const float *normal;           /* (H,W,3) or NULL                       */
    const float *quantile_depths;  /* (H,W,Q) or NULL                       */
    const float *contrib;          /* (N,) or NULL                          */
} pf_raster_grad_in;

/* VisOptions (rasterize.py:15-20).                                         */
typedef struct pf_vis_options {
    float transmittance_threshold;
    int32_t max_intersections;
    float depth_quantile;
    float bkgd_color[3];
} pf_vis_options;

/* Rasterizer.visualize outputs (rasterize.py:2178-2202).                  */
typedef struct pf_vis_out {
    float *color;          /* (H,W,3) */
    float *depth;          /* (H,W)   */
    float *normal;         /* (H,W,3) */
    float *alpha;          /* (H,W)   */
    int32_t *intersections;/* (H,W)   */
} pf_vis_out;

/* Raster faces (prefetch_adjacency_kernel, rasterize.py:580-606): for every
 * CSR edge i->j, faces[e] = (p_j - p_i, r_j) in face_dtype (PF_F16 as the
 * reference).  spheres (N,4) f32.                                          */
int pf_raster_faces(const float *spheres, int64_t num_cells,
                    const int32_t *adjacency, const int32_t *adjacency_offsets,
                    void *faces, int32_t face_dtype, void *stream);

/* Binning pass 1 (count_visible_pinhole_kernel, rasterize.py:140-200):
 * tile_counts[1 + tile] += number of cells whose projected OBB passes the
 * tile SAT.  tile_counts has num_tiles+1 entries and is zeroed by the
 * caller.  cam: one device pf_camera.                                       */
int pf_raster_bin_count(const float *spheres, int64_t num_cells,
                        const pf_camera *cam, int32_t *tile_counts,
                        void *stream);

/* Inclusive scan of n int32 counts into int64 offsets (the reference's
 * torch.cumsum, rasterize.py:1717).                                        */
int pf_raster_scan(const int32_t *counts, int64_t n, int64_t *offsets,
                   void *stream);

/* pf_raster_scan for a list buffer of fixed size, with no host read-back
 * (the reference reads the total to the host, rasterize.py:1718): offsets
 * are clamped to capacity, and a larger total sets PF_STATUS_LIST_CAP in
 * *status (the tiles past the capacity then render incompletely).  Pass
 * num_isect = capacity to pf_raster_bin_write; the whole raster pass is
 * then free of host synchronisation and can be captured in a graph.       */
int pf_raster_scan_bounded(const int32_t *counts, int64_t n, int64_t *offsets,
                           int64_t capacity, int32_t *status, void *stream);

/* Binning pass 2 (write_visible_pinhole_kernel, rasterize.py:293-363) and
 * the per-tile ordering (torch.argsort of the keys, :1771): each tile's
 * cells sorted by power distance from the eye (bit pattern of the positive
 * f32), ties broken by cell index.  write_counter (num_tiles) zeroed by the
 * caller; keys and scratch: num_isect u64 each; tile_prims (num_isect) out;
 * visible_mask (N bytes, caller-zeroed) or NULL gets 1 for listed cells.  */
int pf_raster_bin_write(const float *spheres, int64_t num_cells,
                        const pf_camera *cam, int32_t num_tiles,
                        const int64_t *tile_offsets, int32_t *write_counter,
                        uint64_t *keys, uint64_t *scratch, int32_t *tile_prims,
                        uint8_t *visible_mask, int64_t num_isect, void *stream);

/* forward_kernel (rasterize.py:892-1146): volume objective, f32 attributes,
 * scene.faces = raster faces.  tile_early_stop (num_tiles) out: the list
 * position where the tile's maximum transmittance fell below the
 * threshold (or the list length).  depth_quantiles (H,W,Q) or NULL;
 * ray_gt (H,W,3) needed for point_err / the surface objective.            */
int pf_raster_forward(const pf_scene *scene, const pf_camera *cam,
                      int32_t width, int32_t height,
                      const int64_t *tile_offsets, const int32_t *tile_prims,
                      int32_t *tile_early_stop, float transmittance_threshold,
                      const float *depth_quantiles, int32_t num_quantiles,
                      const float *ray_gt, int32_t return_point_err,
                      int32_t surface_objective, const pf_raster_out *out,
                      void *stream);

/* backward_kernel (rasterize.py:1150-1516): reverse replay from each
 * tile's early-stop position and the saved per-pixel log_t; gradients
 * accumulated into grads (texel_rgb_cam_stride ignored).  K <= 8.          */
int pf_raster_backward(const pf_scene *scene, const pf_camera *cam,
                       int32_t width, int32_t height,
                       const int64_t *tile_offsets, const int32_t *tile_prims,
                       const int32_t *tile_early_stop,
                       const float *depth_quantiles, int32_t num_quantiles,
                       const float *log_t, const pf_raster_grad_in *gin,
                       const pf_grads *grads, void *stream);

/* benchmark_kernel (rasterize.py:614-734): colour only, per-pixel early
 * stop, scene.faces = walk faces (dp, dpm) in f16, attributes f32 or f16.  */
int pf_raster_benchmark(const pf_scene *scene, const pf_camera *cam,
                        int32_t width, int32_t height,
                        const int64_t *tile_offsets, const int32_t *tile_prims,
                        float transmittance_threshold, float *color,
                        void *stream);

/* visualization_kernel (rasterize.py:738-888): scene.faces = raster faces
 * in f16; attributes f32 or f16.                                           */
int pf_raster_visualize(const pf_scene *scene, const pf_camera *cam,
                        int32_t width, int32_t height,
                        const int64_t *tile_offsets, const int32_t *tile_prims,
                        const pf_vis_options *options, const pf_vis_out *out,
                        void *stream);

/* ---- Čech complex (AABBTree) ------------------------------------------- */

/* Grid for the overlap search, in levels: level l has cells of size
 * cell * 2^l from origin and dims ((dims[k] - 1) >> l) + 1; a sphere is
 * listed in the finest level whose cells are at least its diameter, so the
 * last level's cell must be >= 2 r_max.  levels 0 or 1: one uniform grid
 * (cell >= 2 r_max).  Buckets: pf_cech_grid_size of them.  (pf_power_cells
 * reads origin, cell and dims only.)                                      */
typedef struct pf_cech_grid {
    float origin[3];
    float cell;
    int32_t dims[3];
    int32_t levels;
} pf_cech_grid;

/* Bounds of the recovered sphere centres, the largest and the mean radius
 * (AABBTree.update, bvh.py:76-94): bounds[0..2] min, [3..5] max, [6] r_max,
 * [7] mean radius (device floats, 8).  workspace: 7 int32 (device).       */
int pf_cech_bounds(const float *centers, const float *radii, int64_t num_cells,
                   int32_t *workspace, float *bounds, void *stream);

/* Buckets of the grid's levels together (G below), or -1 when the grid is
 * out of range (more than 16 levels, or 2^31 buckets or more).           */
int64_t pf_cech_grid_size(const pf_cech_grid *grid);

/* count_adjacent (bvh.py:7-33) plus its cumsum (:109): buckets the spheres
 * into the grid, counts each sphere's overlapping neighbours and scans the
 * counts into row_offsets (N+1, int64).  grid_counts (G+1) and grid_fill
 * (G) zeroed by the caller; cell_key, grid_items: N; grid_start: G+1;
 * row_counts: N+1 with row_counts[0] = 0.                                  */
int pf_cech_count(const float *centers, const float *radii, int64_t num_cells,
                  const pf_cech_grid *grid, int32_t *grid_counts,
                  int32_t *grid_fill, int32_t *cell_key, int64_t *grid_start,
                  int32_t *grid_items, int32_t *row_counts,
                  int64_t *row_offsets, void *stream);

/* write_adjacent (bvh.py:36-69): fills each CSR row (sorted ascending) and
 * the int32 offsets (N+1).  adjacency holds num_edges = row_offsets[N]
 * entries; workspace: pf_cech_write_workspace_bytes(num_edges, num_cells)
 * bytes, any contents (the unsorted rows and the row sort's storage).
 * -1: sizes out of range.                                                 */
int64_t pf_cech_write_workspace_bytes(int64_t num_edges, int64_t num_cells);
int pf_cech_write(const float *centers, const float *radii, int64_t num_cells,
                  const pf_cech_grid *grid, const int64_t *grid_start,
                  const int32_t *grid_items, const int64_t *row_offsets,
                  int64_t num_edges, int32_t *adjacency,
                  int32_t *adjacency_offsets, void *workspace,
                  int64_t workspace_bytes, void *stream);

/* ---- serving ------------------------------------------------------------ */
/* No reference counterpart: the reference has no serving path (its closest
 * relative is the viewer's one-request render loop, viewer.py:514-592).     */

/* A scene resident on the device: f32 attributes, f16 walk faces, K = 8.
 * sh: (N,16,3) degree-3 coefficients evaluated per texel site inside the
 * walk, or NULL to use scene.texel_rgb.                                    */
/* A served scene: f32 attributes, K = 8 texels, PF_F16 (E,4) walk faces;
 * sh (N,16,3) f32 degree-3 coefficients or NULL (then texel_rgb).        */
typedef struct pf_serve_scene {
    pf_scene scene;
    const float *sh;
    int32_t sh_degree;
    int32_t reserved;
} pf_serve_scene;

/* One request's share of a step: tiles [tile_begin, tile_begin+tile_count)
 * of its image, placed at [step_tile0, step_tile0+tile_count) of the step's
 * tile space (entries sorted by step_tile0).  mode 0: forward (color,
 * log_t); mode 1: also the backward of sum |color - target|^2 into grads.
 * texel_rgb: this request's texel colours (N,K,3) f32 from a colour
 * pre-pass for its camera (pf_sh_forward / pf_sv_forward), or NULL to use
 * the scene's colour source (SH evaluated in the walk, or the scene's own
 * texel_rgb).                                                             */
typedef struct pf_serve_request {
    pf_camera cam;
    int32_t scene;
    int32_t mode;
    int32_t tile_begin;
    int32_t tile_count;
    int64_t step_tile0;
    const int32_t *start_cell;  /* device int holding the start cell, or
                                   NULL to use cam.start_cell            */
    float *color;
    float *log_t;
    const float *target;
    pf_grads grads;
    const float *texel_rgb;
    int64_t bucket0;            /* pf_serve_step_grouped, mode 1: the first
                                   of this entry's num_cells buckets (the
                                   step's mode-1 entries own disjoint,
                                   increasing ranges); unused otherwise  */
} pf_serve_request;

/* One serving step.  scenes, requests and desc are device memory; desc =
 * {num_requests, step_tiles, forward_tiles}: entries with mode 0 come first
 * and cover step tiles [0, forward_tiles), mode-1 entries the rest.  Two
 * launches (forward, forward+backward), each over tile_capacity tiles
 * regardless of desc, so the step can be captured in a graph once and
 * replayed.                                                               */
#define PF_SERVE_SH_IN_WALK   1   /* flag of pf_serve_step: entries with no
                                     texel_rgb of SH scenes evaluate the SH
                                     in the walk (otherwise every entry reads
                                     colours from memory: its own texel_rgb
                                     or the scene's)                        */
/* record: NULL (forward+backward entries replay the walk), or a workspace
 * whose segments hold cap * pf_serve_record_slots(tile_capacity) records
 * (counts unused): each forward+backward ray records its contributing
 * segments there and its backward reads them back; rays with more than cap
 * are replayed.  Steps that may run at the same time need separate
 * workspaces.                                                             */
int64_t pf_serve_record_slots(int64_t tile_capacity);
int pf_serve_step(const pf_serve_scene *scenes,
                  const pf_serve_request *requests, const int64_t *desc,
                  int64_t tile_capacity, float transmittance_threshold,
                  int32_t max_steps, int32_t flags, int32_t *status,
                  const pf_walk_record *record, void *stream);

/* One serving step whose forward+backward entries are differentiated
 * grouped by cell, as pf_walk_backward_grouped does for one walk (no
 * reference counterpart; the backward it replaces per tile is the
 * reference's rasterize.py:1371-1516 accumulation).  desc = {num_requests,
 * step_tiles, forward_tiles, buckets}: mode-1 entries cover at most
 * bwd_tile_capacity tiles and own bucket ranges [bucket0, bucket0 +
 * num_cells of their scene) within [0, buckets), buckets <= max_buckets;
 * a table past either limit skips the mode-1 entries and sets
 * PF_STATUS_LIST_CAP.  Forward tiles run as in pf_serve_step (colours from
 * memory); mode-1 entries must read colours from memory too (their own
 * texel_rgb or the scene's).  record: segments and partials hold
 * cap * 64 * bwd_tile_capacity entries, counts 64 * bwd_tile_capacity
 * ints (cap a multiple of 8); workspace: any contents, at least its fixed
 * parts and one sorted entry; pf_serve_grouped_workspace_bytes gives the
 * size that holds min(cap, 48) entries per slot.  Rays with more than cap
 * contributing segments, or past the entries the workspace holds, are
 * replayed in the same launch.  flags: 0 or PF_BWD_EXACT_TEXEL.  Every launch has a fixed
 * shape (graph capture).  Steps that may run at the same time need
 * separate records and workspaces.                                       */
int64_t pf_serve_grouped_workspace_bytes(int64_t bwd_tile_capacity,
                                         int32_t record_cap,
                                         int64_t max_buckets);
int pf_serve_step_grouped(const pf_serve_scene *scenes,
                          const pf_serve_request *requests,
                          const int64_t *desc, int64_t tile_capacity,
                          int64_t bwd_tile_capacity,
                          float transmittance_threshold, int32_t max_steps,
                          int32_t *status, const pf_walk_record *record,
                          int64_t max_buckets, int32_t flags, void *workspace,
                          int64_t workspace_bytes, void *stream);

/* ---- power-diagram adjacency ----------------------------------------- */

/* The power diagram's adjacency (build_power_adjacency, benchmark.py:25-83,
 * which lifts the sites to 4-D for a Qhull hull) for the part of the
 * diagram inside box = {lo[3], hi[3]}: for each site i, the sites j whose
 * bisector face with i meets the box.  One wavefront per site clips the box by
 * the bisectors of its nearest sites in FP64, then certifies each vertex of
 * the cell with an empty-ball lookup in an occupancy pyramid over the grid
 * buckets.  grid (cell size and dims over the sites' bounds, dims <= 2^20
 * per axis) is chosen by the caller; grid_counts (G+1) and grid_fill (G)
 * zeroed by the caller, cell_key and grid_items N, grid_start G+1;
 * workspace: pf_power_workspace_bytes(grid) bytes (any contents; the
 * occupancy pyramid and the bucket scan's partial sums).
 * lists (N, max_nb) int32 (rows sorted ascending), counts (N), flags (N):
 * bit 0 the cell meets the box boundary (its faces outside the box are
 * those of the diagram of the flagged sites alone), bit 1 capacity
 * overflow (max_nb <= 64; the row is incomplete), bit 2 hidden site
 * (empty cell).  lifted (N): the lift L = |p|^2 - r^2 as the caller
 * computed it (the reference lifts in f32, benchmark.py:29-32); the
 * bisector of i and j is 2 x.(p_j - p_i) <= L_j - L_i.  r_max: largest
 * radius; lift_slack >= |L - (|p|^2 - r^2)| (the lift's rounding).       */
int64_t pf_power_workspace_bytes(const pf_cech_grid *grid);
int pf_power_cells(const float *points, const float *lifted, int64_t num_cells,
                   const pf_cech_grid *grid, int32_t *grid_counts,
                   int32_t *grid_fill, int32_t *cell_key, int64_t *grid_start,
                   int32_t *grid_items, void *workspace,
                   int64_t workspace_bytes, const float *box, float r_max,
                   float lift_slack, int32_t max_nb, int32_t *lists,
                   int32_t *counts, int32_t *flags, void *stream);

/* The faces of the power diagram outside box: for each of the num_hull
 * sites whose cell meets the box boundary (flag bit 0 of pf_power_cells),
 * the other such sites j whose bisector face with it has a vertex outside
 * the box.  Outside the box only those sites own points, so this completes
 * pf_power_cells' adjacency.  Faces farther than far from a site are not
 * seen.  lists (num_hull, max_nb) of site ids, counts and flags as
 * pf_power_cells.                                                        */
int pf_power_hull(const float *points, const float *lifted, const int32_t *hull,
                  int32_t num_hull, const float *box, float far,
                  int32_t max_nb, int32_t *lists, int32_t *counts,
                  int32_t *flags, void *stream);

/* ---- cell order ---------------------------------------------------------- */

/* 63-bit Morton codes of points (N,3) f32 over the box [lower, upper]
 * (host float[3] each): per axis q = (p - lower) / max(upper - lower, 1e-6)
 * * 2^21, clamped to [0, 2^21 - 1] and truncated; the bits of x, y, z
 * interleaved with x highest.  Replaces morton_code_kernel (geometry.py:
 * 18-47); sorting the codes stably gives morton_sort (geometry.py:50-85),
 * the order sort_points (scene.py:197-233) puts cells in.  codes (N) int64. */
int pf_morton_codes(const float *points, int64_t num_points,
                    const float *lower, const float *upper, int64_t *codes,
                    void *stream);

/* ---- colour pre-pass ---------------------------------------------------- */

/* Spherical-harmonics colour per texel site and camera (degree 0..3,
 * coeffs (N,(D+1)^2,3) f32, sites (N,K,3) f32):
 * rgb = max(0, 0.5 + sum_l Y_l(normalize(site - eye)) k_l) with the
 * constants of powerfoam/sh_utils.py:4-74, the value+0.5/clamp of
 * powerfoam/color_fn.py:95-98 and its per-texel-site direction (:71).
 * Replaces the colour pre-pass SphericalVoronoi.forward
 * (color_fn.py:188-246) for SH-coloured scenes.  Output (B,N,K,3) with
 * rgb_cam_stride elements between cameras, rgb_dtype PF_F32|PF_F16.       */
int pf_sh_forward(const float *coeffs, int32_t degree, const float *texel_sites,
                  int64_t num_cells, int32_t num_texel_sites,
                  const pf_camera *cams, int32_t num_cams, void *texel_rgb,
                  int32_t rgb_dtype, int64_t rgb_cam_stride, void *stream);

/* Backward of pf_sh_forward: grad_coeffs += sum over cameras and texels of
 * Y_l * dL/drgb (where the unclamped value is positive).  One thread per
 * cell, no atomics.  grad_coeffs (N,(D+1)^2,3) f32.                        */
int pf_sh_backward(const float *coeffs, int32_t degree, const float *texel_sites,
                   int64_t num_cells, int32_t num_texel_sites,
                   const pf_camera *cams, int32_t num_cams,
                   const float *grad_texel_rgb, int64_t grad_cam_stride,
                   float *grad_coeffs, void *stream);

/* ---- spherical-Voronoi colour pre-pass ----------------------------------
 * Replaces SphericalVoronoi._forward / spherical_voronoi_fwd_kernel
 * (powerfoam/color_fn.py:54-101, 188-246).  Per texel site s (num_sites =
 * N*K rows of texel_sites, f32): dir = normalize(s - eye); dir . forward <
 * fov_cos_cutoff gives value 0.5 and weight sum 0, otherwise value =
 * max(sum w_i v_i / sum w_i + 0.5, 0) with w_i = exp(-temp_i |dir - axis_i|).
 * axes/values are (sv_dof, num_sites, 3), temps (sv_dof, num_sites), all
 * of `dtype` (PF_F32 / PF_F16), as are value_out (num_sites,3) and
 * wsum_out (num_sites).  eye/forward are host float[3].                   */
int pf_sv_forward(const float *texel_sites, int64_t num_sites,
                  const float *eye, const float *forward, float fov_cos_cutoff,
                  const void *axes, const void *values, const void *temps,
                  int32_t sv_dof, int32_t dtype, void *value_out,
                  void *wsum_out, void *stream);

/* Replaces SphericalVoronoi._backward / spherical_voronoi_bwd_kernel
 * (powerfoam/color_fn.py:105-184, 253-315).  Writes (does not accumulate)
 * the gradients of lobes with w_i / sum >= 1e-3 at in-view sites; the
 * caller zeroes grad_axes / grad_values / grad_temps first.  No gradient
 * flows to the sites, as in the reference.                               */
int pf_sv_backward(const float *texel_sites, int64_t num_sites,
                   const float *eye, const float *forward, float fov_cos_cutoff,
                   const void *axes, const void *values, const void *temps,
                   int32_t sv_dof, int32_t dtype, const void *wsum,
                   const void *value, const void *grad_value, void *grad_axes,
                   void *grad_values, void *grad_temps, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PF_HIP_H */
