/*
 * cctag_amd -- C ABI of the MI355X-native CCTag detection hot path.
 *
 * Plain C: pointers, sizes and int error codes (0 = OK, < 0 = error; the
 * library never exits).  One handle per frame size and per calling thread;
 * a handle is not thread-safe.  Host buffers passed in are caller-owned;
 * views returned point into handle-owned host memory that stays valid until
 * the next call on the same handle.
 *
 * Each entry point names the reference interface it replaces
 * (alicevision/CCTag v1.0.3, src/cctag/...).
 */
#ifndef CCTAG_AMD_H
#define CCTAG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCTAGX_OK 0
#define CCTAGX_ERR_ARG (-1)        /* bad argument (null pointer, size, level) */
#define CCTAGX_ERR_HIP (-2)        /* a HIP runtime call failed */
#define CCTAGX_ERR_CAPACITY (-3)   /* a fixed capacity was exceeded */
#define CCTAGX_ERR_STATE (-4)      /* call order violated (e.g. view before front end) */
#define CCTAGX_ERR_NO_DEVICE (-5)  /* no gfx950 device visible */
#define CCTAGX_ERR_UNSUPPORTED (-6)/* frame size or parameter outside what is built */

/* Mirrors cctag::Parameters (Params.hpp:108-199); defaults of Params.hpp:27-66. */
typedef struct cctagx_params {
  float canny_thr_low;                    /* _cannyThrLow 0.01 */
  float canny_thr_high;                   /* _cannyThrHigh 0.04 */
  int32_t dist_search;                    /* _distSearch 30 */
  int32_t thr_gradient_mag_in_vote;       /* _thrGradientMagInVote 2500 */
  float angle_voting;                     /* _angleVoting 0 (must be 0) */
  float ratio_voting;                     /* _ratioVoting 4 */
  float average_vote_min;                 /* _averageVoteMin 0 */
  float thr_median_distance_ellipse;      /* _thrMedianDistanceEllipse 3 */
  int32_t maximum_nb_seeds;               /* _maximumNbSeeds 500 */
  int32_t maximum_nb_candidates_loop_two; /* _maximumNbCandidatesLoopTwo 40 */
  int32_t n_crowns;                       /* _nCrowns 3 */
  int32_t min_points_segment_candidate;   /* _minPointsSegmentCandidate 10 */
  int32_t min_votes_to_select_candidate;  /* _minVotesToSelectCandidate 3 */
  float thresh_robust_estimation_of_outer_ellipse; /* 30 */
  float ellipse_growing_elliptic_hull_width;       /* 2.3 */
  int32_t window_size_on_inner_elliptic_segment;   /* 20 */
  int32_t number_of_multires_layers;               /* 4 */
  int32_t number_of_processed_multires_layers;     /* 4 */
  int32_t n_samples_outer_ellipse;                 /* 150 */
  int32_t num_cuts_in_ident_step;                  /* 22 */
  int32_t num_samples_outer_edge_points_refinement;/* 20 */
  int32_t sample_cut_length;                       /* 100 */
  int32_t imaged_center_n_grid_sample;             /* 5 */
  float imaged_center_neighbour_size;              /* 0.2 */
  float min_ident_proba;                           /* 1e-6 */
  int32_t search_for_another_segment;              /* 1 */
  int32_t do_identification;                       /* 1 */
} cctagx_params;

/* cctag::Parameters::Parameters(nCrowns), Params.cpp:38-76. */
void cctagx_params_default(cctagx_params* p, int n_crowns);

/* ---- file formats (host only, no device needed) --------------------------
 * The Parameters XML archive that cctagDetection's `parameterFile` names
 * (boost::archive::xml_iarchive of Parameters::serialize, Params.hpp:210-248,
 * root element "CCTagsParams", ICCTag.cpp:47-58): every serialized member
 * must be present, as the archive requires; members without a field here
 * (_cutsSelectionTrials, _useLMDif, _writeOutput, _maxEdges, _useCuda,
 * _pinned*) are read and ignored.  A missing file returns CCTAGX_ERR_ARG
 * where the reference prints to stderr and returns (ICCTag.cpp:51-54). */
int cctagx_params_parse_xml(const char* text, cctagx_params* out);
int cctagx_params_read_xml(const char* path, cctagx_params* out);
/* The archive text of *p (the members without a field get the reference's
 * defaults); *len = its length, written to buf (NUL-terminated) when cap
 * allows, else CCTAGX_ERR_CAPACITY (buf == NULL: length only). */
int cctagx_params_write_xml(const cctagx_params* p, char* buf, size_t cap, size_t* len);
/* A marker bank text (CCTagMarkersBank::read, CCTagMarkersBank.cpp:60-78):
 * one row per non-empty line, numbers written as "p/q" or decimals; a line
 * stops at the first text that is neither.  *rows, *cols = the bank's size;
 * the ratios are written row-major when ratios != NULL and the bank fits
 * max_rows x max_cols (else CCTAGX_ERR_CAPACITY).  Rows of different lengths:
 * CCTAGX_ERR_UNSUPPORTED.  An unreadable file: CCTAGX_ERR_ARG (the reference
 * throws std::ios_base::failure). */
int cctagx_bank_parse(const char* text, float* ratios, int max_rows, int max_cols, int* rows,
                      int* cols);
int cctagx_bank_read(const char* path, float* ratios, int max_rows, int max_cols, int* rows,
                     int* cols);

typedef struct cctagx_handle cctagx_handle;

/* Replaces TagPipe::initialize(w, h) (cuda/tag.cu:76-127) and the per-pipe
 * allocation of initCuda (Detection.cpp:732-760).  Allocates every device and
 * pinned host buffer for frames of width x height once.  Pyramid level l is
 * (width >> l) x (height >> l), each level resampled from the one above
 * (ImagePyramid.cpp:23-33, Level.cpp:67); a size at which a level would be
 * empty is CCTAGX_ERR_UNSUPPORTED.  The handle's resolution is fixed: every entry point that takes a frame checks its
 * width and height against it and returns CCTAGX_ERR_ARG on a mismatch,
 * where the reference stops the process (Detection.cpp:753-757). */
int cctagx_create(const cctagx_params* params, int width, int height, cctagx_handle** out);
int cctagx_destroy(cctagx_handle* h);

/* The marker bank identification reads: rows x cols radius ratios
 * (R_outer / r_j, innermost ring first, row-major), the table of
 * CCTagMarkersBank (CCTagMarkersBank.cpp:22-78) that cctagDetection's
 * `pBank` / `cctagBankFilename` arguments supply (ICCTag.cpp:60-67,86-93).
 * A handle starts with the built-in bank of params->n_crowns
 * (CCTagMarkersBank.cpp:120-282).  1 <= rows <= 128, 1 <= cols <= 8, every
 * ratio finite and > 0; a bank whose rows differ in length is not
 * supported (CCTAGX_ERR_UNSUPPORTED). */
int cctagx_set_bank(cctagx_handle* h, const float* ratios, int rows, int cols);

/* ---- front end (pyramid .. vote) ---------------------------------------
 * Replaces TagPipe::load + TagPipe::tagframe (cuda/tag.cu:154-216): upload a
 * row-major 8-bit gray frame (stride in bytes) and run every pyramid level
 * through derivatives, Canny, thinning, edge compaction, gradient descent,
 * vote and seed selection on the GPU. */
int cctagx_front_end(cctagx_handle* h, const uint8_t* gray, int width, int height,
                     size_t stride);

/* Host view of one level after the front end; what TagPipe::convertToHost
 * (cuda/frame_export.cu:23-165) and TagPipe::getPlane/getDx/getDy/getEdges
 * export.  All arrays are in raster / point-index order. */
typedef struct cctagx_level_view {
  int32_t width, height;          /* level size */
  int32_t shape_w, shape_h;       /* EdgePointCollection shape (level-0 size) */
  const uint8_t* img;             /* width*height */
  const int16_t* dx;              /* width*height */
  const int16_t* dy;              /* width*height */
  const uint8_t* edges;           /* width*height, thinned Canny map {0,255} */
  const int32_t* edge_map;        /* width*height, point index or -1 */
  int32_t n_edges;
  const int16_t* edge_x;          /* n_edges, raster order */
  const int16_t* edge_y;
  const int16_t* edge_dx;
  const int16_t* edge_dy;
  const int32_t* before;          /* n_edges, descent along -gradient, -1 = none */
  const int32_t* after;           /* n_edges, descent along +gradient */
  const int32_t* chosen;          /* n_edges, voted point, -1 = none */
  const float* total_dist;        /* n_edges, field-line length of the vote */
  const int32_t* voters_index;    /* n_edges + 1, CSR offsets */
  const int32_t* voters_list;     /* voters_index[n_edges], ascending per point */
  const float* flow_length;       /* n_edges, running mean of total_dist over voters */
  const int32_t* is_max;          /* n_edges, vote count when >= min votes, else -1 */
  int32_t n_seeds;
  const int32_t* seeds;           /* n_seeds, in the order they reached min votes */
} cctagx_level_view;

int cctagx_get_level(cctagx_handle* h, int level, cctagx_level_view* out);

/* ---- full detection ------------------------------------------------------
 * One marker as cctag::ICCTag exposes it (ICCTag.hpp:28-82): imaged centre
 * x(), y(), id(), getStatus() (1 = identified; negative codes of
 * CCTag.hpp:416-426), rescaledOuterEllipse(), plus quality, pyramid level,
 * scale and the homography. */
typedef struct cctagx_marker {
  float x, y;
  int32_t id;
  int32_t status;
  float quality;
  int32_t pyramid_level;
  float scale;
  float ellipse[5];      /* rescaled outer ellipse: cx, cy, a, b, angle */
  float homography[9];   /* row-major, marker plane -> image (CCTag::homography) */
} cctagx_marker;

/* Replaces cctag::cctagDetection(markers, pipeId, frame, graySrc, params,
 * durations, bank) (ICCTag.hpp:124-131, Detection.cpp:773-999): front end and
 * identification on the GPU, candidate stage on the host.  Writes up to
 * max_out markers sorted by id and their total count in *n_out (which may
 * exceed max_out).  `frame` is informational, as in the reference. */
int cctagx_detect(cctagx_handle* h, const uint8_t* gray, int width, int height, size_t stride,
                  uint64_t frame, cctagx_marker* out, int max_out, int* n_out);

/* The markers of the last cctagx_detect / cctagx_detect_loaded (up to
 * max_out; *n_out = their count), for a caller whose buffer was too small. */
int cctagx_last_markers(cctagx_handle* h, cctagx_marker* out, int max_out, int* n_out);

/* The two halves of cctagx_detect.  cctagx_load is TagPipe::load
 * (cuda/tag.cu:154-159): the frame is copied into device memory and stays
 * resident.  cctagx_detect_loaded runs the detection on the resident frame
 * (TagPipe::tagframe onwards), so a caller can time or repeat detection
 * without the host-to-device copy. */
int cctagx_load(cctagx_handle* h, const uint8_t* gray, int width, int height, size_t stride);
int cctagx_detect_loaded(cctagx_handle* h, uint64_t frame, cctagx_marker* out, int max_out,
                         int* n_out);

/* Color input.  The reference's applications convert every image and video
 * frame with cv::cvtColor(src, gray, COLOR_BGR2GRAY) before cctagDetection
 * (applications/detection/main.cpp:357,427,515; regression/TestLog.cpp:93,
 * 114); these take the 8-bit BGR (channels 3) or BGRA (channels 4, alpha
 * ignored) frame instead and convert it on the device with OpenCV's integer
 * weights, gray = (1868 B + 9617 G + 4899 R + 8192) >> 14.  stride is in
 * bytes (>= width * channels).  The first call on a handle allocates a
 * device staging buffer of width * height * channels bytes.  Otherwise as
 * cctagx_load / cctagx_detect. */
int cctagx_load_bgr(cctagx_handle* h, const uint8_t* bgr, int width, int height, size_t stride,
                    int channels);
int cctagx_detect_bgr(cctagx_handle* h, const uint8_t* bgr, int width, int height, size_t stride,
                      int channels, uint64_t frame, cctagx_marker* out, int max_out, int* n_out);

/* ---- identification of given markers ------------------------------------
 * The identification loop of cctagDetection (Detection.cpp:873-965:
 * identify_step_1, then TagPipe::uploadCuts / imageCenterOptLoop /
 * imageCenterRetrieve and identify_step_2, Identification.cpp:1172-1480)
 * for markers the caller supplies, on the frame of the last
 * cctagx_front_end, cctagx_detect or cctagx_detect_loaded (the frame must
 * have been through the front end: CCTAGX_ERR_STATE otherwise).  Markers are
 * identified in list order, as the reference does; the conditioning scale of
 * the grid search is that of the first marker refined in this call. */
typedef struct cctagx_ident_marker {
  float ellipse[5];     /* rescaledOuterEllipse: cx, cy, a, b, angle (level-0 pixels) */
  float center[2];      /* initial imaged centre (CCTag::centerImg) */
  float scale;          /* CCTag::scale(): 2^pyramid level of the detection */
  int32_t n_points;     /* rescaledOuterEllipsePoints */
  const float* points;  /* n_points x 4: x, y, gx, gy */
} cctagx_ident_marker;
typedef struct cctagx_ident_result {
  int32_t status;       /* 1 identified, else CCTag.hpp:416-426 */
  int32_t id;           /* -1 when no ID was read */
  float x, y;           /* imaged centre after the grid search */
  float quality;        /* 1 / normalised residual (CCTag::setQuality) */
  float homography[9];  /* row-major, marker plane -> image */
  int32_t n_cuts;       /* cuts identify_step_1 selected (<= 32) */
  float stops[32][2];   /* their stops: refined outer-edge points */
} cctagx_ident_result;
int cctagx_identify(cctagx_handle* h, const cctagx_ident_marker* in, int n,
                    cctagx_ident_result* out);

/* Host threads of the candidate stage, the counterpart of the reference's
 * tbb::parallel_for loops (Detection.cpp:616,644,680).  A handle starts with
 * min(16, hardware threads) or $CCTAGX_HOST_THREADS; 0 restores that default,
 * 1 runs the stage on the calling thread.  Results do not depend on it. */
int cctagx_set_host_threads(cctagx_handle* h, int threads);
int cctagx_get_host_threads(cctagx_handle* h);

/* Centre refinement of the identification (no counterpart in the reference,
 * which always runs the grid search of refineConicFamilyGlob,
 * Identification.cpp:813-1158): 0, the default, runs that grid search;
 * steps > 0 runs its first iteration and then `steps` descent steps on the
 * differentiable score (cctagx_refiner_refine) and reads the ID out at the
 * centre reached.  Centres then differ from the grid search's by a fraction
 * of a pixel and an ID can in rare cases differ. */
int cctagx_set_refinement(cctagx_handle* h, int gradient_steps);

/* Breakdown of the last cctagx_detect: host wall time after the front end
 * (staging, candidate stage, identification control) and device time of the
 * two identification kernels, microseconds.  The reference's logtime::Mgmt
 * probes (Detection.cpp:789-964) report the same stage boundaries. */
int cctagx_detect_times(cctagx_handle* h, float* host_us, float* select_us, float* refine_us);

/* Where the time and work of the last detection went: device time of the
 * front end (first to last kernel), host wall time of staging the level
 * outputs and of the host part after it, device time of the two
 * identification kernels, the markers each received and the bilinear image
 * samples each evaluated.  Counterpart of the reference's logtime::Mgmt
 * stage probes (Detection.cpp:789-964). */
typedef struct cctagx_detect_stats {
  float front_end_us;
  float stage_us;
  float host_us;
  float select_us;
  float refine_us;
  int32_t n_select;
  int32_t n_refine;
  int64_t select_samples;
  int64_t refine_samples;
  /* algorithmic FLOPs of the grid search: per marker, iteration and grid
   * point 25 per cut sample and 3 per sample of every cut pair */
  int64_t refine_flops;
  /* host wall time of the stages inside host_us: edge sets, seed linking
   * (loop one), flow components (loop two), segment labels, candidate
   * checks (loop three), level-0 re-fit, identification (including its two
   * kernels) and de-duplication */
  float load_us, loop_one_us, loop_two_us, labels_us, loop_three_us, rescale_us, ident_us,
      finish_us;
  /* inside ident_us: host preparation and result reading, and the two
   * identification steps (transfers and kernels).  The cut selection and
   * the refinement are one backend call, counted in ident_refine_us;
   * ident_select_us is zero */
  float ident_prep_us, ident_select_us, ident_refine_us, ident_post_us;
  /* bytes of candidate-stage input copied to the host, and 1 when the frame
   * was staged in two steps (its counts read back, then the gather) because
   * its levels did not fit the buffer of the one-step gather, which then
   * grows for the next frame */
  int64_t staged_bytes;
  int32_t staged_two_step;
} cctagx_detect_stats;
int cctagx_get_detect_stats(cctagx_handle* h, cctagx_detect_stats* out);

/* ---- batched detection ---------------------------------------------------
 * Many frames of uneven sizes detected together: the front end of every
 * frame and level runs as one captured hipGraph launch over a single plane
 * set, the candidate stage of the frames runs in parallel on the host, and
 * identification runs one launch per step for the markers of all frames.
 * Each frame's markers are exactly what cctagx_detect returns for it.  No
 * counterpart in the reference, whose pipes take one frame per call
 * (Detection.cpp:732-760). */
typedef struct cctagx_batch cctagx_batch;
int cctagx_batch_create(const cctagx_params* params, int n_frames, const int32_t* widths,
                        const int32_t* heights, cctagx_batch** out);
int cctagx_batch_destroy(cctagx_batch* b);
/* Copies frame `frame` (row-major 8-bit, stride in bytes) to the device;
 * width and height must be those given at cctagx_batch_create. */
int cctagx_batch_load(cctagx_batch* b, int frame, const uint8_t* gray, int width, int height,
                      size_t stride);
/* Work partition of the batched front end: filter tile variant (0, 1, 2 =
 * 64x16, 64x32, 64x64 tiles) and blocks per plane of the edge-indexed
 * kernels.  -1 / 0 leave the choice to the autotuner, which times the
 * candidates on the first detection's frames before capturing the graph. */
int cctagx_batch_set_partition(cctagx_batch* b, int filter_variant, int edge_blocks);
/* cctagx_set_refinement for the batch. */
int cctagx_batch_set_refinement(cctagx_batch* b, int gradient_steps);
/* Detects every loaded frame.  Writes up to max_out markers, frame by frame
 * (each frame sorted by id); frame_first[f] .. frame_first[f+1] (n_frames + 1
 * entries) index frame f's markers, and *n_out is the total, which may
 * exceed max_out. */
int cctagx_batch_detect(cctagx_batch* b, cctagx_marker* out, int max_out, int32_t* frame_first,
                        int* n_out);
typedef struct cctagx_batch_stats {
  float front_end_us;   /* device time of the graph launch */
  float stage_us;       /* host wall time of staging every plane */
  float candidates_us;  /* host wall time of all frames' candidate stages */
  float ident_us;       /* host wall time of the batched identification */
  float select_us, refine_us; /* device time of the two identification kernels */
  float tune_us;        /* autotuning time spent in this call (0 after the first) */
  int32_t filter_variant, edge_blocks, graph;
  int32_t n_markers;
} cctagx_batch_stats;
int cctagx_batch_get_stats(cctagx_batch* b, cctagx_batch_stats* out);

/* ---- differentiable centre refinement ------------------------------------
 * The identification's centre search as a differentiable function: for a
 * marker with outer ellipse E and cut stops p_k, the score of a homography H
 * (marker plane -> image) is the mean over valid cut pairs of the summed
 * squared differences of their half-bilinear samples taken along
 * pi(H (t q_k, 1)), q_k = pi(H^-1 p_k) -- what refineConicFamilyGlob
 * (Identification.cpp:813-1158) evaluates at every grid point.  A forward /
 * adjoint kernel pair returns the score and its gradient; a batched descent
 * moves every marker's imaged centre c along -dC/dc with H = H(c) the closed
 * form of computeHomographyFromEllipseAndImagedCenter (:739-798); the
 * reference's grid search runs on the same markers as the fallback and as
 * the result to compare with.  No counterpart in the reference. */
typedef struct cctagx_refiner cctagx_refiner;
typedef struct cctagx_refine_marker {
  float ellipse[5];     /* outer ellipse: cx, cy, a, b, angle (image coordinates) */
  float center0[2];     /* initial imaged centre */
  int32_t n_cuts;       /* 2 .. 32 */
  float stops[32][2];   /* cut stops on the outer edge */
} cctagx_refine_marker;
int cctagx_refiner_create(const cctagx_params* params, int max_markers, cctagx_refiner** out);
int cctagx_refiner_destroy(cctagx_refiner* r);
/* The image every marker is sampled in (row-major 8-bit, stride in bytes). */
int cctagx_refiner_set_image(cctagx_refiner* r, const uint8_t* gray, int width, int height,
                             size_t stride);
int cctagx_refiner_set_markers(cctagx_refiner* r, const cctagx_refine_marker* markers, int n);
/* Forward + adjoint at given homographies (n x 9, row-major): score (n) and
 * gradient (n x 10: d/dh00 .. d/dh21 with h22 held, then d/d of a
 * marker-plane translation (x, y) of the centre, H' = H [1 0 x; 0 1 y; 0 0 1]).
 * fp64 != 0 samples in double (checking), else in float.  A marker with
 * fewer than two valid cuts scores DBL_MAX with a zero gradient. */
int cctagx_refiner_cost_grad(cctagx_refiner* r, const float* H, int fp64, double* cost,
                             double* grad);
/* H(c) of the closed form at given centres (n x 2) -> n x 9. */
int cctagx_refiner_homography(cctagx_refiner* r, const float* centers, float* H);
/* `steps` batched descent steps from center0: centres (n x 2) and scores (n). */
int cctagx_refiner_descend(cctagx_refiner* r, int steps, int fp64, float* centers, double* cost);
/* The grid search's first iteration (its 5 x 5 neighbourhood of center0,
 * the centre it moves to), then `steps` descent steps from there: centres
 * (n x 2) and scores (n).  The coarse step keeps the descent in the right
 * basin when center0 is several pixels off. */
int cctagx_refiner_refine(cctagx_refiner* r, int steps, int fp64, float* centers, double* cost);
/* The reference grid search on the same markers: centres (n x 2) and ok flags (n). */
int cctagx_refiner_grid(cctagx_refiner* r, float* centers, int32_t* ok);
/* Device time of the last descend (or refine) and grid calls, microseconds. */
int cctagx_refiner_times(cctagx_refiner* r, float* descend_us, float* grid_us);

/* No counterpart in the reference: an entry point for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  cctagx_refiner_set_image + cctagx_refiner_set_markers
 * + cctagx_refiner_grid in one synchronous call that allocates nothing: the
 * frame (host memory, at most CCTAGX_STATIC_MAX_PIXELS pixels), its quad
 * image, up to CCTAGX_STATIC_MAX_MARKERS markers' jobs and the grid search's
 * state go into one fixed device array of the library (24.3 MB, held by every
 * GPU the library runs on until the process ends); larger inputs return
 * CCTAGX_ERR_CAPACITY.  Runs on the current device's null stream and writes
 * the same centres (n x 2) and ok flags (n) as cctagx_refiner_grid.  All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU
 * race and must be serialized by the caller. */
#define CCTAGX_STATIC_MAX_PIXELS (1920 * 1088)
#define CCTAGX_STATIC_MAX_MARKERS 64
int cctagx_refine_grid_static(const cctagx_params* params, const uint8_t* gray, int width,
                              int height, size_t stride, const cctagx_refine_marker* markers,
                              int n, float* centers, int32_t* ok);

/* ---- timing ------------------------------------------------------------
 * Stage times of the last front end in microseconds, measured with HIP
 * events on the library stream; names are static strings.  By default the
 * front end is timed as one stage ("seeds": start to end); per_kernel != 0
 * records an event after every kernel (each record is a packet on the
 * stream, so per-kernel timing is for profiling passes). */
int cctagx_set_stage_timing(cctagx_handle* h, int per_kernel);
int cctagx_stage_count(cctagx_handle* h);
int cctagx_stage_time(cctagx_handle* h, int i, const char** name, float* usec);
/* Times `iters` back-to-back front ends on the resident frame (no upload)
 * between two HIP events and returns the average in microseconds. */
int cctagx_time_front_end(cctagx_handle* h, int iters, float* usec_per_frame);

/* Library / device info.  cctagx_last_error() describes the last failure
 * of the calling thread. */
const char* cctagx_version(void);
const char* cctagx_last_error(void);
int cctagx_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* CCTAG_AMD_H */
