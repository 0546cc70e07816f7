/*
 * dplus_mi355x.h -- C ABI of libdplus_mi355x, the MI355X-native scattering
 * hot path of D+ (amplitude grids, symmetry docking, orientation averages).
 *
 * Plain C types only.  Every function returns 0 on success or an error code
 * of the reference's ErrorCode enum (Common/Common.h:184-228); the message of
 * the last failure on the calling thread is available from
 * dplus_last_error() as JSON {"code": int, "message": str}, the convention
 * PythonBackendWrapper::ConvertException uses (Backend/Backend/PythonBackendWrapper.cpp:49-60).
 *
 * All device work runs on one GPU (device 0 unless DPLUS_DEVICE is set).
 */
#ifndef DPLUS_MI355X_H
#define DPLUS_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DPLUS_PRECISION_FP64 0  /* FP64 phases, sincos and accumulation */
#define DPLUS_PRECISION_MIXED 1 /* FP32 bulk, FP64 range reduction and OA accumulation */

#define DPLUS_OA_MC 0    /* "Monte Carlo (Mersenne Twister)" */
#define DPLUS_OA_VEGAS 1 /* "Adaptive (VEGAS) Monte Carlo" */
#define DPLUS_OA_GK 2    /* "Adaptive Gauss Kronrod" */

#define DPLUS_CALC_ATOMIC_FORMFACTORS 1 /* Backend/GPU/atomicFormDefines.h:5 */
#define DPLUS_CALC_DUMMY_SOLVENT 2      /* Backend/GPU/atomicFormDefines.h:6 */

/* Last error of the calling thread as JSON; valid until the next call. */
const char *dplus_last_error(void);

/* Library/device description as JSON into buf (truncated to len). */
int dplus_device_info(char *buf, int len);

/* ------------------------------------------------------------------ grid */
/* Opaque Jacobian spherical grid: device-resident node array plus an
 * optional host mirror in the reference layout (2*P doubles, complex
 * interleaved, Backend/Backend/Grid.cpp:1078-1111).
 * Replaces JacobianSphereGrid(gridSize, qMax) (Backend/Backend/Grid.h:231-327,
 * bound by PythonInterface/dplus/wrappers/Grid.pxd:20-48). */
typedef struct dplus_grid dplus_grid;

/* JacobianSphereGrid::JacobianSphereGrid(u16, double) (Grid.cpp:204-211);
 * an odd grid_size is bumped by one as Amplitude does (Amplitude.cpp:1628). */
int dplus_grid_create(int grid_size, double qmax, int precision, dplus_grid **out);
void dplus_grid_destroy(dplus_grid *g);
/* GetSize, GetQMax, stepSize, GetActualSize (= shells N), point count, precision */
int dplus_grid_info(const dplus_grid *g, int *grid_size, double *qmax, double *step,
                    int *actual_size, long long *points, int *precision);
/* GetDataPointer (Grid.cpp:757, Grid.pyx:50-56): host values, 2*P doubles.
 * The host mirror reflects the device after dplus_grid_download(). */
double *dplus_grid_data(dplus_grid *g);
/* GetInterpolantPointer (Grid.cpp:1207): host spline coefficients, 2*P doubles */
double *dplus_grid_interp(dplus_grid *g);
/* IndexFromIndices / IndicesFromIndex (Grid.cpp:1078-1111) */
long long dplus_grid_index_from_indices(int qi, long long ti, long long pi);
void dplus_grid_indices_from_index(long long index, int *qi, long long *ti, long long *pi);
/* host values -> device, then CalculateSplines (Grid.cpp:1113-1151) on the GPU */
int dplus_grid_upload(dplus_grid *g);
/* CalculateSplines on the device values */
int dplus_grid_calculate_splines(dplus_grid *g);
/* device values and spline coefficients -> host mirror */
int dplus_grid_download(dplus_grid *g);
/* GetSphr(q, theta, phi) (Grid.cpp:487-548) for n points; out = 2n doubles */
int dplus_grid_get_sphr(dplus_grid *g, int n, const double *q, const double *theta,
                        const double *phi, double *out);
/* InterpolateThetaPhiPlane(ri, theta, phi) (Grid.cpp:61-134, bound by
 * Grid.pyx:139-140) for n points: the phi-Hermite / theta four-point
 * interpolation inside shell ri[i] (1 <= ri <= actual size), theta in [0, pi],
 * phi in [0, 2 pi]; out = 2n doubles */
int dplus_grid_interpolate_theta_phi_plane(dplus_grid *g, int n, const int *ri, const double *theta,
                                           const double *phi, double *out);
/* CalculateIntensity(q, [theta,] eps, seed, iterations, phi_min, phi_max)
 * (Grid.cpp:393-484) with the reference's std::mt19937 stream; theta < 0 draws
 * theta at random and, like the reference's overload, ignores the phi range. */
int dplus_grid_calculate_intensity(dplus_grid *g, double q, double theta, double eps,
                                   unsigned seed, unsigned long long iterations, double phi_min,
                                   double phi_max, double *result);

/* ------------------------------------------------------------------ PDB leaf */
/* Atoms and solvent description of PDBAmplitude (Backend/Backend/PDBAmplitude.cpp:1199-1336):
 * ion[i] indexes the nion rows of coef9 (a1,b1,a2,b2,a3,b3,a4,b4,c; X-ray table,
 * PDBReaderLib/PDBReaderLib.cpp:2260-2293) and rad (nm).  Boxes (dims in voxels
 * of side vox_step, centre in nm) come in two sets: the outer-solvent layer
 * added with box_contrast (rho_out - rho_0, PDBAmplitude.cpp:298-314) and the
 * excluded volume of radius types 1, 2, 3 and 5 added with excl_contrast
 * (-rho_0: the solvent it displaces, PDBAmplitude.cpp:277-296). */
typedef struct {
    int natoms;
    const double *xyz; /* 3*natoms, nm */
    const int *ion;    /* natoms */
    int nion;
    const double *coef9; /* 9*nion */
    const double *rad;   /* nion */
    double sol_ed;       /* rho_0 for the dummy-atom solvent, e/nm^3 */
    double c1;
    int flags; /* DPLUS_CALC_* */
    double scale;
    int nbox;
    const int *box_dims;  /* 3*nbox */
    const double *box_com; /* 3*nbox */
    double vox_step;
    double box_contrast;
    int nexcl;
    const int *excl_dims;  /* 3*nexcl */
    const double *excl_com; /* 3*nexcl */
    double excl_contrast;
} dplus_pdb_desc;

/* PDBAmplitude::calculateGrid (PDBAmplitude.cpp:319-435, GPU path
 * PDBJacobianGridAmplitudeCalculation PDBAmplitudeCUDA.cu:203-436) followed by
 * CalculateSplines; host_copy != 0 also fills the host mirror. */
int dplus_pdb_build_grid(dplus_grid *g, const dplus_pdb_desc *d, int host_copy);

/* Device-resident copy of a PDB leaf's atoms and solvent boxes, sorted and
 * tiled once (the state PDBAmplitude::PreCalculate prepares,
 * PDBAmplitude.cpp:1199-1336), so repeated grid builds read HBM only. */
typedef struct dplus_structure dplus_structure;
int dplus_structure_create(const dplus_pdb_desc *d, dplus_structure **out);
void dplus_structure_destroy(dplus_structure *s);
/* build the amplitude grid of a resident structure (+ CalculateSplines) */
int dplus_structure_build_grid(dplus_structure *s, dplus_grid *g, int host_copy);
/* Grid i of n from structure i (+ CalculateSplines), the mixed-precision ones
 * as grouped builder launches of up to 16 grids each (structures, grid sizes
 * and atom counts may all differ), the same values as n single builds.  The
 * grids must be distinct (ERROR_INVALID_ARGS otherwise).  The reference
 * builds one leaf at a time (Amplitude.cpp:1348-1405). */
int dplus_structure_build_grids(int n, dplus_structure *const *s, dplus_grid *const *g, int host_copy);

/* Autotuner of the mixed builder's launch shape (no counterpart in the
 * reference, whose kernels have fixed launch shapes, PDBAmplitudeCUDA.cu:
 * 242-247, 304-399): times every workgroup size (64, 128, 256, 512 units) with every
 * atom-slice count (1..8) on this structure and grid, records the fastest in
 * the process's shape table under (grid size, tile-count class) and builds the
 * grid (+ splines).  Later builds of that class use the recorded shape (in the
 * request engine only with its autotune flag: shapes timed on an idle device
 * do not carry over to a device shared by concurrent requests); others use
 * the occupancy model.  Writes the chosen shape, its time and the
 * occupancy model's (ms). */
int dplus_structure_autotune(dplus_structure *s, dplus_grid *g, int *wg, int *slices, double *best_ms,
                             double *model_ms);
/* the shape table as text, one "grid_shells tile_class wg slices" line per
 * entry (callee-allocated, free with dplus_free); loading such text (lines
 * starting with '#' are comments) adds or replaces entries; clearing empties
 * the table */
/* the launch shape a mixed build of grid g from structure s takes, and whether
 * it comes from the shape table (1) or the occupancy model (0) */
int dplus_structure_build_shape(dplus_structure *s, dplus_grid *g, int *wg, int *slices, int *tuned);
int dplus_build_shapes_save(char **text, long long *len);
int dplus_build_shapes_load(const char *text);
int dplus_build_shapes_clear(void);
/* The form of every later FP64 grid build in this process (no counterpart in
 * the reference, whose FP64 path is its CPU calcAmplitude, PDBAmplitude.cpp:243-317):
 * 0 (default) the paired builder's algebra in double -- Friedel halving, node
 * pairs at phi and phi + pi sharing the sincos of q_x x + q_y y, the row's
 * sincos of q_z z shared through LDS, atom slices; 1 the plain form, one node
 * at a time with one double sincos per (node, source) term.  Both give the
 * direct sum's values up to rounding order.  The environment variable
 * DPLUS_FP64_PLAIN selects 1 at load. */
int dplus_set_fp64_builder(int plain);

/* ------------------------------------------------------------------ measurement */
/* Per-kernel HIP-event timing on the library's stream: enable/disable, read the
 * launch count and summed duration (ms) of a kernel family ("build", "splines",
 * "mc", "vegas", "gk", "dock"), reset. */
int dplus_timing_enable(int on);
int dplus_timing_get(const char *kernel, long long *launches, double *ms);
int dplus_timing_reset(void);
/* Throughput microbenchmark of gfx950 VALU ops: ops/s of v_fma_f32, v_sin_f32,
 * v_cos_f32 and v_fma_f64 over the whole device (for the transcendental weight). */
int dplus_microbench_valu(double *fma32, double *sin32, double *cos32, double *fma64);
/* A plain device copy of `bytes` (multiple of 16) timed in ms: HBM bandwidth probe and the
 * known-size kernel hardware byte counters are calibrated against. */
int dplus_copy_probe(long long bytes, double *ms);

/* Gather ceiling of the orientation averages (no counterpart in the reference):
 * the node loads of a mixed MC average without its arithmetic -- for each of
 * `directions` stratified directions (a multiple of 256) of every layer in
 * [layer0, layer0 + nlayers) and each of nrot rotations (alpha, beta, gamma,
 * scale per rotation; nrot = 0: the drawn direction itself), 4 planes x 4 theta
 * rows x 2 phi nodes of 16 B -- at full occupancy on a mixed grid.  Writes the
 * launch time (ms, after one untimed launch) and the lookups it made (each
 * 512 B of nodes). */
int dplus_gather_probe(dplus_grid *g, int nrot, const double *rot, long long directions, int layer0, int nlayers,
                       double *ms, double *lookups);

/* ------------------------------------------------------------------ orientation average */
/* A leaf of the hybrid regime (Amplitude.cpp:2028-2303): a grid seen under nrot
 * rotations rot[4r..4r+3] = (alpha, beta, gamma [rad], scale), rotation r placed
 * at ntrans[r] translations (3 doubles each, nm, packed in rotation order). */
typedef struct {
    dplus_grid *grid;
    int nrot;
    const double *rot;
    const int *ntrans;
    const double *trans;
} dplus_leaf;

typedef struct {
    int method;             /* DPLUS_OA_* */
    unsigned long long seed;
    long long max_dirs;     /* MC: directions per shell group ("OrientationIterations") */
    int chunk;              /* MC: directions between convergence checks and the stratification period (multiple of 256) */
    double eps;             /* "Convergence"; <= 0 runs the full count */
    int vegas_iterations;
    int vegas_points_per_bin; /* 32x32 bins; points per iteration = 1024 * this */
    long long gk_max_depth;
} dplus_oa_params;

typedef struct {
    long long dirs_used;  /* MC directions per group actually used (max over groups) */
    int iterations_used;  /* VEGAS iterations (max over groups) */
    long long evaluations; /* GK integrand evaluations */
    double seconds;       /* wall time of the call */
} dplus_oa_stats;

/* DomainModel::CalculateIntensityVector OA step: GPUCalcMCOAJacobSphrDD
 * (Backend/GPU/OAJacob.cu:132-383), GPUHybridCalculator::ComputeIntensity
 * (Backend/GPU/GPUHybridCalc.cu:164-720) and GaussKron2DSphereRecurs
 * (Backend/Backend/Amplitude.cpp:1416-1467).  q must be ascending. */
int dplus_orientation_average(int nleaf, const dplus_leaf *leaves, int nq, const double *q,
                              const dplus_oa_params *p, double *intensity, dplus_oa_stats *st);

/* Allocation-free variant of dplus_orientation_average, with no counterpart in
 * the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Monte
 * Carlo or VEGAS only (ERROR_INVALID_ARGS for Gauss-Kronrod); the same kernels
 * and the same intensities as dplus_orientation_average.  Every device buffer
 * (leaf tables, q and shell groups, MC block partials, VEGAS bin sums and
 * history) comes from one static 64 MiB device array per GPU, held until the
 * process ends; a call that would need more is rejected with
 * ERROR_INVALID_ARGS before any work.  Calls on one GPU share the array, so
 * overlapping calls would race on it: the library serialises them. */
int dplus_orientation_average_noalloc(int nleaf, const dplus_leaf *leaves, int nq, const double *q,
                                      const dplus_oa_params *p, double *intensity, dplus_oa_stats *st);

/* Azimuthal average of the 2-D generate path (DomainModel::CalculateIntensity2DMatrix,
 * Backend/Backend/Amplitude.cpp:843-993, its per-point MC over phi in
 * AverageIntensitiesBetweenLayers2D :2825-2897): intensity[i] = mean over nphi
 * stratified azimuths phi of |A(q[i], theta[i], phi)|^2 for the docked leaves,
 * q[i] in [0, qmax], theta[i] in [0, pi]; phi from a Philox stream keyed by
 * (seed, sample, point). */
int dplus_azimuthal_average(int nleaf, const dplus_leaf *leaves, int npts, const double *q, const double *theta,
                            int nphi, unsigned long long seed, double *intensity);

/* ------------------------------------------------------------------ manual symmetry */
/* ManualSymmetry::calculateGrid (DefaultModels/Symmetries.cpp:879-962, GPU
 * ConstructManualSymmetryJacobianKernel Backend/GPU/ManSymGenJacob.cu:35-282):
 * out(q) = scale * sum over leaves/rotations/translations of A_leaf(R^T q) e^{i q.t},
 * evaluated at every node of out, then CalculateSplines. */
int dplus_dock_grid(dplus_grid *out, int nleaf, const dplus_leaf *leaves, double scale);

/* ------------------------------------------------------------------ space-filling symmetry */
/* GridSymmetry::calculateGrid (DefaultModels/Symmetries.cpp:348-431, GPU
 * CalcSpaceFillingKernelJcb Backend/GPU/SymmetryGenJacob.cu:32-132):
 * out(q) = scale * [sum over leaves/rotations/translations of A_leaf(R^T q) e^{i q.t}]
 *          * prod_k sum_{n<N_k} e^{i n q.v_k},
 * with the basis v_k built from distances[3] (nm), angles_deg[3] (alpha, beta,
 * gamma) and reps[3] as GridSymmetry::PreCalculate does (Symmetries.cpp:59-97),
 * including its argument checks (ERROR_INVALIDARGS). */
int dplus_dock_grid_space_filling(dplus_grid *out, int nleaf, const dplus_leaf *leaves, double scale,
                                  const double *distances, const double *angles_deg, const double *reps);

/* ------------------------------------------------------------------ hydration layer */
/* Voxelised hydration ("outer solvent") layer of PDBAmplitude::CalculateSolventSpace
 * (Backend/Backend/VoxelizedSolvent.cpp:4-194, box reduction :263-363) with the
 * reference's voxel marking: the part of the probe shell (atom radius + probe)
 * that faces the bulk solvent, grown into the bulk to the solvation thickness;
 * cavities count as atoms when fill_holes != 0.  Output: boxes (dims in voxels,
 * centres in nm).  Writes up to `capacity` boxes and sets *nbox to the number
 * found; call again with capacity >= *nbox when it is larger. */
int dplus_hydration_boxes(int n, const double *xyz, const double *rad, double probe, double thickness,
                          double vox, int fill_holes, int capacity, int *nbox, int *dims, double *com);
/* The excluded volume of the same voxel colouring (the "atoms" voxels that
 * ReduceSolventSpaceToIrregularBoxes turns into solventBoxDims/COM,
 * VoxelizedSolvent.cpp:4-194,263-362): voxels whose corner lies within an atom
 * radius, cavities when fill_holes != 0, and -- with probe > 0 -- the part of the
 * probe shell that the hydration layer does not take.  Radius types 1, 2, 3 and
 * 5 subtract it with the solvent density (PDBAmplitude.cpp:277-296).  Same
 * capacity protocol as dplus_hydration_boxes. */
int dplus_excluded_volume_boxes(int n, const double *xyz, const double *rad, double probe, double thickness,
                                double vox, int fill_holes, int capacity, int *nbox, int *dims, double *com);

/* ------------------------------------------------------------------ scripted symmetry */
/* LuaSymmetry::PreCalculate (Backend/Backend/Symmetry.cpp:463-521) with the
 * checks of VerifyAndObtainData (:579-605): runs a symmetry script (Lua 5.1
 * text; the library carries its own interpreter for the subset such scripts
 * use, without os and io) and calls Populate(p, nlayers) with p[i][j] =
 * params[(i - 1) * nlp + j - 1] (1-based, j up to the script's Information.NLP,
 * missing entries 0).  The result's rows are {x, y, z, alpha, beta, gamma}
 * (nm, degrees), written row-major to rows[6 * k + c] when *nrows <= capacity;
 * call again with capacity >= *nrows when it is larger.  A script error is
 * ERROR_INVALIDARGS (21) with the script's message. */
int dplus_lua_symmetry_rows(const char *script, int nlayers, int nlp, const double *params, int capacity, int *nrows,
                            double *rows);

/* ------------------------------------------------------------------ amplitude files */
/* AmpGridAmplitude::ReadAmplitudeFromFile (Backend/Backend/Amplitude.cpp:3724-3796):
 * a ".amp" legacy stream (JacobianSphereGrid::ImportBinaryData, Grid.cpp:272-345)
 * or an ".ampj" zip with stored members grid.dat / criticalinfo.json /
 * header.json (PythonInterface/dplus/Amplitudes.py:705-750), uploaded to a new
 * grid of the given precision and splined.  *header (optional) receives the
 * file's header text; release it with dplus_free. */
int dplus_amp_read(const char *path, int precision, dplus_grid **out, char **header);
/* Amplitude::WriteAmplitudeToFile (Amplitude.cpp:3119-3190) for ".ampj" paths,
 * JacobianSphereGrid::ExportBinaryData (Grid.cpp:847-894) for ".amp" paths. */
int dplus_amp_write(dplus_grid *g, const char *path, const char *header);

/* ------------------------------------------------------------------ backend */
/* The job interface of PythonBackendWrapper (Backend/Backend/PythonBackendWrapper.h:15-46,
 * bound by PythonInterface/dplus/wrappers/BackendWrapper.pxd), for the model
 * types of the scattering hot path: "PDB" leaves, "AMP" leaves (amplitude
 * files), Manual Symmetry (",26"), Space-filling Symmetry (",25") and
 * Scripted Symmetry (a Lua script's rows) nodes, orientation averaged by Monte
 * Carlo, VEGAS or adaptive Gauss-Kronrod.  Strings returned through char**
 * are malloc'ed; release them with dplus_free. */
typedef struct dplus_backend dplus_backend;
void dplus_free(void *p);
int dplus_backend_create(dplus_backend **out);
void dplus_backend_destroy(dplus_backend *b);
/* CheckCapabilities (PythonBackendWrapper.cpp:20-31): ERROR_NO_GPU (31) without a device */
int dplus_backend_check_capabilities(dplus_backend *b, int check_tdr);
/* GetAllModelMetadata (PythonBackendWrapper.cpp:33-47): the supported models, JSON */
int dplus_backend_get_all_model_metadata(dplus_backend *b, char **out);
/* InitializeCache(cacheDir) (PythonBackendWrapper.cpp:76-85) */
int dplus_backend_initialize_cache(dplus_backend *b, const char *dir);
/* StartGenerate(state, useGPU) (PythonBackendWrapper.cpp:87-106, BackendWrapper.cpp:313-372):
 * state = CalculationInput args JSON {"state": {...}, "x": [q...], ...}.
 * ERROR_ILLEGAL_JSON (9) for unparsable text, ERROR_INVALIDPARAMTREE (19) for a
 * malformed tree, ERROR_JOBRUNNING (2) while a job runs; the job itself runs
 * on a worker thread. */
int dplus_backend_start_generate(dplus_backend *b, const char *state, int use_gpu);
/* StartGenerate2D(state, useGPU) (PythonBackendWrapper.cpp:108-128, BackendWrapper.cpp:376-420):
 * the same state and job protocol; the job computes the (q_z, q_perp) intensity
 * matrix over every pair of x values (DomainModel::CalculateIntensity2DMatrix,
 * Amplitude.cpp:843-993) as azimuthal averages (dplus_azimuthal_average). */
int dplus_backend_start_generate_2d(dplus_backend *b, const char *state, int use_gpu);
/* GetGenerate2DResults (PythonBackendWrapper.cpp:160-174, BackendWrapper.cpp:583-624):
 * {"Headers": [...], "2DGraph": [[I(q_z = x[i], q_perp = x[j]) for j] for i]} */
int dplus_backend_get_generate_2d_results(dplus_backend *b, char **out);
/* GetJobStatus (PythonBackendWrapper.cpp:130-143, Conversions.cpp:49-64):
 * {"isRunning": bool, "progress": double, "code": int (-1 while running), "message": str} */
int dplus_backend_get_job_status(dplus_backend *b, char **out);
/* GetGenerateResults (PythonBackendWrapper.cpp:145-158, BackendWrapper.cpp:540-581):
 * {"Headers": [{"ModelPtr": n, "Header": str}...], "Graph": [I(q)...]} */
int dplus_backend_get_generate_results(dplus_backend *b, char **out);
/* GetModelPtrs (PythonBackendWrapper.cpp:213-223): pointers of the last state's
 * models; writes min(*n, capacity) entries and sets *n to the total */
int dplus_backend_get_model_ptrs(dplus_backend *b, unsigned *ptrs, int capacity, int *n);
/* GetAmplitude (PythonBackendWrapper.cpp:188-198): the model's grid as a legacy
 * .amp byte stream of *len bytes */
int dplus_backend_get_amplitude(dplus_backend *b, unsigned model_ptr, char **bytes, long long *len);
/* GetPDB (PythonBackendWrapper.cpp:200-211, LocalBackend.cpp:969-1015): the
 * atoms of a PDB model in its own frame, or of a symmetry assembled over its
 * sub-locations after REMARK lines with its header, as 80-byte ATOM records
 * (PDBAmplitude.cpp:1469-1538, Symmetry.cpp:234-351); *len bytes (records can
 * hold zero bytes past their text, as the reference's do), released with
 * dplus_free.  ERROR_JOBNOTFOUND (3) before any job, ERROR_MODELNOTFOUND (14)
 * for a model without atoms. */
int dplus_backend_get_pdb(dplus_backend *b, unsigned model_ptr, char **text, long long *len);
/* SaveAmplitude (PythonBackendWrapper.cpp:176-186): .ampj (or .amp) file */
int dplus_backend_save_amplitude(dplus_backend *b, unsigned model_ptr, const char *path);
/* Stop (PythonBackendWrapper.cpp:225-235): the running job ends with ERROR_STOPPED (1) */
int dplus_backend_stop(dplus_backend *b);

/* ------------------------------------------------------------------ request engine */
/* On-device serving of I(q) requests; no counterpart in the reference, whose
 * backend runs one job per process (JobManager.cpp:137-205).  Requests are
 * admitted in arrival order as device slots (one HIP stream each) and the HBM
 * budget allow, run concurrently, and give the same result alone or in a mix. */
typedef struct dplus_engine dplus_engine;

typedef struct {
    long long hbm_budget_bytes; /* admission budget; <= 0: 85 % of device memory */
    int streams;                /* concurrent requests (1..32), one worker thread and HIP
                                   stream each; <= 0: 16 */
    int batch_builds;           /* 1: mixed PDB builds of at most batch_terms terms that are
                                   ready together run as one grouped builder launch (up to
                                   16 grids of different structures and sizes) */
    long long batch_terms;      /* canonical nodes x (atoms + boxes); <= 0: 1e10 */
    int autotune;               /* 1: builds take the shape table's shapes, and a mixed build
                                   whose (grid, tile class) the table lacks is autotuned on
                                   first sight (dplus_structure_autotune); 0: the occupancy
                                   model's shapes */
    int graphs;                 /* 1: a fixed-count request (Monte Carlo or VEGAS with eps <= 0)
                                   runs its kernel chain (builder, splines, orientation
                                   average, chi^2, result download) as one captured HIP graph,
                                   updating an instantiation its slot keeps for the chain's
                                   shape; the same bits as the eager launches */
} dplus_engine_config;

typedef struct {
    const dplus_pdb_desc *pdb; /* a PDB leaf built for this request, or NULL with `grid` */
    dplus_grid *grid;          /* a resident leaf grid (outlives the request) instead of `pdb` */
    int grid_size;             /* grid of the PDB leaf */
    double qmax;
    int precision;             /* DPLUS_PRECISION_* of the PDB leaf */
    int nrot;                  /* placements as in dplus_leaf; 0: the leaf once at the origin */
    const double *rot;
    const int *ntrans;
    const double *trans;
    int nq;
    const double *q;           /* ascending */
    dplus_oa_params oa;
    const double *data;        /* optional: chi^2 of fit_scale * I + fit_constant against data/sigma */
    const double *sigma;
    double fit_scale;
    double fit_constant;
} dplus_request;

typedef struct {
    double queue_ms;    /* submission to start on the device */
    double run_ms;      /* on the device */
    double latency_ms;  /* submission to result */
    double chi2;
    int chi2_on_device; /* 1 when chi^2 was reduced on the device */
    int slot;           /* stream the request ran on */
    long long reserved_bytes;
    dplus_oa_stats oa;
    int graph;          /* 1 when the request ran as a captured HIP graph */
} dplus_request_result;

int dplus_engine_create(const dplus_engine_config *cfg, dplus_engine **out);
/* waits for every admitted and queued request, then releases the engine */
void dplus_engine_destroy(dplus_engine *e);
/* copies the request's inputs and queues it; ERROR_OUT_OF_MEMORY (13) when it
 * could never fit the budget */
int dplus_engine_submit(dplus_engine *e, const dplus_request *r, long long *ticket);
int dplus_engine_poll(dplus_engine *e, long long ticket, int *done);
/* blocks until the request is done, writes nq intensities and releases the ticket */
int dplus_engine_wait(dplus_engine *e, long long ticket, double *intensity, dplus_request_result *res);
int dplus_engine_stats(dplus_engine *e, long long *submitted, long long *completed, long long *peak_reserved,
                       long long *budget, long long *max_concurrent);

/* structure cache of the engine (no counterpart in the reference): requests
 * whose PDB description has the same content (128-bit hash of every array and
 * scalar) share one device-resident structure; at most budget/16 (<= 8 GiB) is
 * cached, and admission keeps that much of the budget free of requests */
int dplus_engine_cache_stats(dplus_engine *e, long long *hits, long long *misses, long long *bytes,
                             long long *capacity);
/* grouped builder launches of the engine (no counterpart in the reference):
 * launches made and builds they carried */
int dplus_engine_batch_stats(dplus_engine *e, long long *launches, long long *builds);
/* captured requests of the engine (no counterpart in the reference): graph
 * launches, instantiations and in-place updates of kept instantiations, and
 * the host milliseconds summed over requests spent capturing (waits for the
 * process-wide capture lock included) and updating or instantiating */
int dplus_engine_graph_stats(dplus_engine *e, long long *launches, long long *instantiations, long long *updates,
                             double *capture_ms, double *exec_ms);

/* ------------------------------------------------------------------ fit sessions */
/* Fitting-loop evaluation over resident subunit grids (the hybrid docking of
 * Amplitude.cpp:2028-2303 with DomainModel's MC average, OAJacob.cu:132-383),
 * replayed as one HIP graph per placement shape; chi^2 against data/sigma is
 * reduced on the device.  No counterpart in the reference's API. */
typedef struct dplus_session dplus_session;
/* leaves: resident grids (same size, qmax, precision); fixed q, data, sigma,
 * seed and direction count (rounded up to a multiple of 256).  The directions
 * are stratified in chunks of 8192 when that divides the count (the same
 * directions as dplus_orientation_average with chunk 8192), else as one chunk. */
int dplus_session_create(int nleaf, dplus_grid *const *leaves, int nq, const double *q, const double *data,
                         const double *sigma, unsigned long long seed, long long directions,
                         dplus_session **out);
/* placements[l] places leaf l (its grid field may be NULL); writes nq
 * intensities and chi^2 of fit_scale * I + fit_constant */
int dplus_session_evaluate(dplus_session *s, const dplus_leaf *placements, int nleaf, double fit_scale,
                           double fit_constant, double *intensity, double *chi2);
int dplus_session_info(dplus_session *s, long long *evaluations, long long *graph_captures);
void dplus_session_destroy(dplus_session *s);

/* ------------------------------------------------------------------ synthetic inputs */
/* Seeded self-avoiding random walk (step, minimum separation in nm) confined to
 * the ellipsoid with semi-axes (ax, ay, az); xyz receives 3*n coordinates. */
int dplus_synth_walk(unsigned long long seed, int n, double step, double min_sep, double ax,
                     double ay, double az, double *xyz);

#ifdef __cplusplus
}
#endif
#endif /* DPLUS_MI355X_H */
