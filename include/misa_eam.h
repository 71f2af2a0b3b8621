/*
 * misa_eam.h - C-ABI of the MI355X EAM step library (libmisa_eam_hip.so).
 *
 * The reference (MISA-MD v0.4.0) reaches an accelerator through eight C++
 * hooks named ${ARCH_NAME}_* (src/arch/arch_imp.h:17-31), selected at build
 * time by MD_HIP_ARCH_ENABLE_FLAG (config.cmake:17, arch/arch_configure.cmake
 * :13-16) and linked into libmd as target md_arch_hip
 * (arch/arch_libs.cmake:17-38).  Their arguments are C++ types from libcomm
 * (comm::BccDomain), pot (eam) and MISA-MD itself (NeighbourIndex,
 * AtomElement).  This header flattens them into plain C: every argument is a
 * pointer plus sizes, every call returns 0 on success or a negative MISA_E_*
 * code, and misa_last_error() gives the message.  A thin C++ adapter that
 * defines the eight hip_* symbols on top of these calls makes the library a
 * drop-in md_arch_hip at one rank.
 *
 * Two modes share one handle:
 *  - hook mode (misa_lat_rho / misa_lat_df / misa_lat_force) works on the
 *    reference's own host array of 104-byte AtomElement slots, synchronously,
 *    exactly where atom::latRho/latDf/latForce call the hooks;
 *  - resident mode (misa_upload_state ... misa_step ... misa_download_state)
 *    keeps the whole state in HBM and runs complete velocity-Verlet steps
 *    (lattice and inter atoms, ownership decisions, periodic halo) on the
 *    GPU; it replaces the loop body of simulation::simulate
 *    (simulation.cpp:164-194), for which the reference has no hook.
 *
 * Threading: one handle per host thread; calls on a handle are serialised
 * and run on that handle's HIP stream.  Every call returns after its work
 * has finished on the device.
 */
#ifndef MISA_EAM_H
#define MISA_EAM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct misa_handle misa_handle;
typedef struct misa_batch misa_batch;

/* error codes */
#define MISA_OK 0
#define MISA_E_ARG -1        /* bad argument or call order */
#define MISA_E_HIP -2        /* HIP runtime error */
#define MISA_E_IO -3         /* potential file missing or malformed */
#define MISA_E_DEVICE -4     /* a kernel raised a device-side error flag */
#define MISA_E_CAPACITY -5   /* a fixed-capacity device list overflowed */

/* precision of the resident-mode force path */
#define MISA_PREC_FP64 0     /* FP64 everywhere */
#define MISA_PREC_MIXED 1    /* FP32 splines and pair terms, FP64 positions/Verlet/sums */
#define MISA_PREC_MIXED_KAHAN 17   /* MIXED with per-atom sums in FP32 plus Kahan compensation (the lattice
                                      force terms joining it in chunks of eight offsets) */
#define MISA_PREC_MIXED_F32SUM 33  /* MIXED with plain FP32 per-atom sums */

/* Geometry of the single-rank domain: the whole periodic box.  Mirrors the
 * comm::BccDomain fields the reference reads (simulation.cpp:29-56):
 * phase_space (cells per dimension), lattice_const, cutoff_radius_factor;
 * the ghost width is ceil(cutoff_radius_factor) + 1 cells as the reference
 * sets it (simulation.cpp:46). */
typedef struct misa_geom {
  int64_t phase_space[3];
  double lattice_const;
  double cutoff_radius_factor;
} misa_geom;

/* Thermodynamic reductions over the resident state (FP64 on the device). */
typedef struct misa_thermo {
  double mvv;            /* sum m v^2 over lattice and inter atoms (system_configuration.cpp:61-84) */
  double kinetic;        /* mvv / 2 (reduceEnergy, system_configuration.cpp:113-126) */
  double temperature;    /* mvv * mvv2e / ((3N-3) k_B) (system_configuration.cpp:51-59) */
  double e_embed;        /* sum_i F_i(rho_i) */
  double e_pair;         /* 1/2 sum_{i != j} phi_ij(r_ij) */
  double e_potential;    /* e_embed + e_pair */
  double virial;         /* sum over pairs of r_ij . f_ij */
  double force_sum[3];   /* sum_i f_i (system_configuration.cpp:8-24) */
  double max_speed;      /* max_i |v_i| */
  int64_t n_lattice;     /* occupied lattice sites */
  int64_t n_inter;       /* inter atoms in the local list */
  int64_t n_vacancy;     /* vacant lattice sites */
} misa_thermo;

/* ---- set-up: hip_env_init / hip_domain_init (arch_env.hpp:15-19,
 *      simulation.cpp:52-54) ---- */

/* Create a handle on HIP device `device` for the domain `g`. */
int misa_create(const misa_geom* g, int device, misa_handle** out);
/* hip_env_clean (arch_env.hpp:21-25): free every device buffer of `h`. */
int misa_destroy(misa_handle* h);
/* Message of the last failing call on `h` (never NULL). */
const char* misa_last_error(const misa_handle* h);

/* hip_nei_offset_init (simulation.cpp:67-69, neighbour_index.h:22-25): the
 * four linear-offset lists NeighbourIndex::make builds in the ghost-extended
 * doubled-x array (neighbour_index.inl:13-76): full lists for even and odd x
 * and the lexicographically positive halves. */
int misa_set_offsets(misa_handle* h, const int64_t* even, int64_t n_even, const int64_t* odd, int64_t n_odd,
                     const int64_t* half_even, int64_t n_half_even, const int64_t* half_odd, int64_t n_half_odd);
/* Build the same four lists on the host, as NeighbourIndex::make does. */
int misa_make_offsets(misa_handle* h);
/* Copy list `which` (0 even, 1 odd, 2 half even, 3 half odd) into `out`;
 * returns its length through `n`. */
int misa_get_offsets(const misa_handle* h, int which, int64_t* out, int64_t cap, int64_t* n);

/* hip_pot_init (simulation.cpp:133): parse a DYNAMO setfl (.eam.alloy) file
 * and build the interpolation tables the pot library builds in
 * eam::interpolateFile (simulation.cpp:131). */
int misa_set_potential_file(misa_handle* h, const char* setfl_path);
/* Copy the 7-coefficient table (per knot: derivative cubic c0..c2, value
 * cubic c3..c6) of kind 0 = F(rho), 1 = rho(r), 2 = r*phi(r) for file element
 * index `idx` (pair index i*(i+1)/2+j for kind 2).  Returns its length. */
int misa_get_potential_table(const misa_handle* h, int kind, int idx, double* out, int64_t cap, int64_t* n);

/* ---- hook mode: hip_eam_rho_calc / hip_eam_df_calc / hip_eam_force_calc
 *      (atom.cpp:159-160, 293-294, 320-321) ----
 * `atoms` is the host AtomList array (104-byte AtomElement slots, ghost
 * extended, atom_list.cpp:18).  The caller has zeroed rho and f
 * (atom::clearForce).  On return, rho (resp. df, f) of every non-vacant
 * sub-box slot holds the complete lattice-lattice sum; ghost slots keep
 * zero partials, so the reference's reverse halos add nothing. */
int misa_lat_rho(misa_handle* h, void* atoms, double cutoff_radius);
int misa_lat_df(misa_handle* h, void* atoms, double cutoff_radius);
int misa_lat_force(misa_handle* h, void* atoms, double cutoff_radius);
/* The same three hooks (which = 0 rho, 1 df, 2 force) without any device
 * allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc), since misa_lat_rho/df/force stage the
 * array in a device buffer that their first call allocates.  This variant
 * stages it in a fixed 32 MiB array that the library holds on every GPU it
 * runs on until the process ends, so boxes of more than 32 MiB / 104 B =
 * 322,638 ghost-extended slots (about 48^3 cells) are refused with
 * MISA_E_ARG.  All handles on one GPU share that array: calls that overlap
 * on one GPU (from several host threads) race.  Each call overwrites the
 * whole staged array, so nothing carries over between calls. */
int misa_lat_hook_noalloc(misa_handle* h, int which, void* atoms, double cutoff_radius);

/* ---- resident mode ---- */

/* Upload the whole state: the ghost-extended AtomElement array (only the
 * sub-box slots are read) and the inter-atom list (n_inter AtomElement
 * records in list order). */
int misa_upload_state(misa_handle* h, const void* atoms, const void* inter, int64_t n_inter);
/* Download into the same layout; ghost slots are filled from their periodic
 * images as exchangeAtom leaves them.  `n_inter` returns the inter count;
 * MISA_E_CAPACITY if it exceeds `inter_cap`. */
int misa_download_state(misa_handle* h, void* atoms, void* inter, int64_t inter_cap, int64_t* n_inter);
/* Build the initial state the way WorldBuilder::build does
 * (world_builder.cpp:64-103) and upload it: BCC sites, velocities from
 * std::mt19937(seed), alloy types from glibc rand() after srand(alloy_seed)
 * (alloy_seed = 1 reproduces the reference's unseeded rand()), zero
 * momentum, rescale to tset when tset != 0. */
int misa_build_world(misa_handle* h, uint32_t seed, double tset, const int32_t ratio[3], uint32_t alloy_seed);

int misa_set_precision(misa_handle* h, int mode);
/* NewtonMotion::setTimestepLength (newton_motion.cpp:13-17). */
int misa_set_timestep(misa_handle* h, double dt);
/* Adaptive step length (no counterpart in the reference, which only switches
 * dt per stage, md_simulation.cpp:42-58): with max_disp > 0 every step of
 * misa_step takes dt = min(dt_max, max_disp / max_i |v_i|) over lattice and
 * inter atoms, computed on the device from the velocities at the start of the
 * step; max_disp = 0 returns to the fixed step of misa_set_timestep. */
int misa_set_adaptive_dt(misa_handle* h, double dt_max, double max_disp);
/* Neighbour skin of the density pass, both precisions (Angstrom; default 0.2, 0 turns
 * it off; no counterpart in the reference, whose lists are rebuilt every
 * step).  A full scan of the pruned offsets also records, per wavefront, the
 * offsets within cutoff + skin; until the lattice atoms' summed per-step
 * displacement reaches skin / 2 (or a refill, upload or setter change), the
 * density pass visits only those.  The pairs and their order do not change,
 * so neither do the results. */
int misa_set_neighbour_skin(misa_handle* h, double skin);
/* Ownership events since the last upload (atom::decide, atom.cpp:21-84):
 * [0] lattice atoms taken off their sites (moved more than 0.2a), [1] inter
 * atoms put back on a vacant site of their Wigner-Seitz cell, [2] of those
 * the ones farther than 0.2a from it (their neighbourhoods then visit the
 * whole full neighbour list for one step). */
int misa_defect_counters(misa_handle* h, int64_t out[3]);
/* Simulated time (ps) since the last upload, and the length of the last step. */
int misa_sim_time(misa_handle* h, double* t_ps, double* dt_last);
/* Inter-atom list capacity (default 65536 entries). */
int misa_set_inter_capacity(misa_handle* h, int64_t cap);
/* Device error flags.  Every call that runs kernels checks them when it
 * ends: a NaN or infinite force or electron density (MISA_E_DEVICE), an inter
 * atom outside the ghost-extended box (MISA_E_DEVICE), a full inter-atom list
 * (MISA_E_CAPACITY).  An electron density beyond the last knot of the F(rho)
 * table (F' then comes from the table's last segment, as the interpolation
 * clamps it) is recorded always and fails the call only with the range check
 * on (on = 1; default off).  The flags stay set until the next upload.  No
 * counterpart in the reference, whose hooks have no error channel
 * (arch_imp.h:17-31); its development build aborts on |sum f_x| > 1e-5
 * (md_simulation.cpp:150-164) and on a wrong F' halo length
 * (df_embed_packer.cpp:50-53). */
#define MISA_FLAG_INTER_CAPACITY 1
#define MISA_FLAG_GHOST_CAPACITY 2
#define MISA_FLAG_NON_FINITE 4
#define MISA_FLAG_LEFT_BOX 8
#define MISA_FLAG_RHO_RANGE 16
int misa_set_range_check(misa_handle* h, int32_t on);
/* The flags word as the last checking call read it (MISA_FLAG_* bits). */
int misa_device_flags(const misa_handle* h, int64_t* flags);

/* clearForce + computeEam on the resident state (simulation.cpp:141-143):
 * the first force evaluation before the step loop. */
int misa_compute_forces(misa_handle* h);
/* `nsteps` iterations of simulation::simulate's loop body
 * (simulation.cpp:164-194): half kick and drift, decide, inter-atom
 * migration and periodic halo, EAM forces, half kick. */
int misa_step(misa_handle* h, int64_t nsteps);
/* atom::setv (atom.cpp:475-494): add the PKA velocity to the atom at lattice
 * coordinate (lat[0], lat[1], lat[2], basis lat[3]). */
int misa_setv(misa_handle* h, const int32_t lat[4], const double dir[3], double energy);
/* simulation::collisionStep (simulation.cpp:208-217): setv, halo, forces. */
int misa_collision_step(misa_handle* h, const int32_t lat[4], const double dir[3], double energy);
/* configuration::rescale (system_configuration.cpp:86-111). */
int misa_rescale(misa_handle* h, double tset);
/* Reductions over the current state (forces from the last evaluation). */
int misa_thermo_compute(misa_handle* h, misa_thermo* out);

/* Device time (ms, from HIP events on the handle's stream) of the most
 * recent misa_step call, split by kernel class: [0] rho/df pass, [1] force +
 * Verlet pass, [2] inter-atom and ownership work, [3] whole call. */
int misa_last_timings(const misa_handle* h, double out[4]);
/* Record HIP events around every kernel launch of misa_step (on = 1) so
 * misa_last_timings splits the device time by kernel class;
 * misa_last_launch_counts gives the number of launches per class. */
int misa_set_profiling(misa_handle* h, int on);
int misa_last_launch_counts(const misa_handle* h, int64_t out[3]);

/* ---- batches of independent boxes (no counterpart in the reference, which
 *      runs one box per MPI job) ----
 * A batch steps many configured handles of one device together (replicas of
 * other sizes, temperatures, compositions or cascades).  Each step's launches
 * of every box are captured once into a hipGraph (two loop bodies per graph,
 * every box an independent branch) and replayed, so the boxes' kernels
 * overlap on the GPU and cost no per-launch host work.  The graph is
 * re-captured when a box changes (precision, step, offsets, potential,
 * capacity, tile shape).  The handles stay owned by the caller. */
int misa_batch_create(misa_handle* const* handles, int n, misa_batch** out);
int misa_batch_destroy(misa_batch* b);
const char* misa_batch_last_error(const misa_batch* b);
/* `nsteps` loop bodies of every box, as misa_step would run them one box at
 * a time; `graph_replays` (may be NULL) returns the replays so far. */
int misa_batch_step(misa_batch* b, int64_t nsteps, int64_t* graph_replays);
/* Autotuner of the site kernels' launch shape: times `steps` steps for every
 * y-tile height dividing the box (up to 32) and, in mixed precision, every
 * count of neighbour gathers in flight per wave (4 or 8 in the density pass,
 * 2 or 4 in the force pass), and keeps the fastest; the trial steps advance
 * the simulation.  The arithmetic of a site does not depend on the shape. */
int misa_autotune(misa_handle* h, int64_t steps, int32_t* best_ty, double* best_ms);
/* Set / read the launch shape directly (y-tile rows, gathers in flight). */
int misa_set_launch_shape(misa_handle* h, int32_t ty, int32_t gathers_rho, int32_t gathers_force);
int misa_get_launch_shape(const misa_handle* h, int32_t* ty, int32_t* gathers_rho, int32_t* gathers_force);
/* Force passes (mixed and FP64) in persistent blocks that hold the majority
 * species' spline knots in LDS (the mixed pass also the pruned offsets):
 * MISA_LDS_AUTO (default; boxes large enough that every block walks several
 * virtual blocks), MISA_LDS_ON, MISA_LDS_OFF.  No counterpart in the
 * reference (a tuning switch of this build's kernels); the pairs, their
 * order and their rounding are the same either way, so results are
 * bit-identical. */
#define MISA_LDS_AUTO 0
#define MISA_LDS_ON 1
#define MISA_LDS_OFF 2
int misa_set_lds_tables(misa_handle* h, int32_t mode);
/* Work-balanced XCD slabs (default on; no counterpart in the reference, a
 * placement choice of this build's kernels).  The site passes give each of
 * the device's 8 XCDs a z-slab of the box; with balancing on, the first
 * mixed LDS force pass of every misa_step call measures the cycles spent per
 * z-plane and, after the call, the slabs are re-cut so that each holds about
 * an eighth of that work (a cascade's disordered core otherwise keeps one
 * XCD busy while the others wait).  Off returns to the even cut.  Every site
 * sees the same pairs in the same order either way, so the trajectory does
 * not change. */
int misa_set_xcd_balance(misa_handle* h, int32_t on);
/* The slab bounds: XCD c walks planes [zlo[c], zlo[c+1]) of the box. */
int misa_xcd_slabs(const misa_handle* h, int32_t zlo[9]);
/* Set the slab bounds (zlo[0] = 0, zlo[8] = planes of the box,
 * non-decreasing, no slab above twice an eighth of the box's sites; an empty
 * slab leaves its XCD idle) and turn balancing off. */
int misa_set_xcd_slabs(misa_handle* h, const int32_t zlo[9]);
/* The force pass the last force evaluation launched (-1 before any):
 * kForce (FP64), kForceM (mixed, spline segments from global memory),
 * kForceL (mixed, persistent blocks with the majority species' segments in
 * LDS) or kForceDL (FP64, persistent blocks with the majority species'
 * segments in LDS).  The LDS passes run where MISA_LDS_* picks them; the
 * library falls back to kForceM / kForce when the device refuses their LDS
 * size. */
#define MISA_FORCE_PASS_FP64 0
#define MISA_FORCE_PASS_MIXED 1
#define MISA_FORCE_PASS_MIXED_LDS 2
#define MISA_FORCE_PASS_FP64_LDS 3
int misa_force_pass_used(const misa_handle* h, int32_t* kind);

/* ---- host mirror of libcomm's domain at one rank ----
 *
 * comm::BccDomain (built as simulation::createDomain does,
 * simulation.cpp:28-50, ghost size ceil(f) + 1) for the rank that owns the
 * whole periodic box; the headers misa-md_amd/host/comm/{domain,comm}.h hold
 * the C++ API (Region, BccDomain::Builder, fwCommLocalRegion, Packer,
 * neiSendReceive) and misa-md_amd/host/logs/logs.h the kiwi::logs calls.
 * Regions are {x_low, y_low, z_low, x_high, y_high, z_high}. */
typedef struct misa_domain {
  int64_t phase_space[3];
  double lattice_const;
  double cutoff_radius_factor;
  int64_t cut_lattice;
  int32_t grid_size[3];
  int32_t grid_coord[3];
  int32_t rank_id_neighbours[6]; /* [dimension][direction] */
  int32_t pad_;
  double meas_global_length[3];
  double meas_global_region[6];
  double meas_sub_box_region[6];
  int64_t sub_box_lattice_size[3];
  int64_t lattice_size_ghost[3];
  int64_t ghost_extended_lattice_size[3];
  int64_t sub_box_lattice_region[6];
  int64_t ghost_ext_lattice_region[6];
  int64_t dbx_sub_box_lattice_size[3];
  int64_t dbx_lattice_size_ghost[3];
  int64_t dbx_ghost_extended_lattice_size[3];
  int64_t dbx_sub_box_lattice_region[6];
  int64_t dbx_ghost_ext_lattice_region[6];
} misa_domain;
int misa_host_domain(const misa_geom* g, misa_domain* out);
/* comm::fwCommLocalRegion: the slab of the ghost-extended array (x doubled,
 * indices from 0) sent to the neighbour in (dimension, direction). */
int misa_host_fw_region(const misa_geom* g, int dimension, int direction, int64_t region[6]);
/* AtomList::exchangeAtomFirst (atom_list.cpp:25-49) at one rank on a host
 * AtomElement array: every ghost slot gets the type and the shifted position
 * of its periodic image (LatPackerFirst, lat_particle_packer.cpp:22-160). */
int misa_host_exchange_atoms(const misa_geom* g, void* atoms);

/* ---- frontend I/O: atom dumps (frontend/io), one rank ----
 *
 * Copy mode (AtomDump, frontend/io/atom_dump.cpp:44-101, with
 * BufferedFileWriter, buffered_io.cpp:17-49): 72-byte AtomInfoDump records
 * (atom_info_dump.h:14-22: id, step, type, inter_type 0, x, v) after a
 * 128-byte file header (zeros) and a 128-byte rank header whose first 8
 * bytes hold the record count, in the layout tools/dump2text/out_convert.cpp
 * reads.  A frame is the inter atoms in list order, then the occupied
 * sub-box sites of the ghost-extended AtomElement array in z, y, doubled-x
 * order (atom_dump.cpp:66-84).  Frames append; close writes the count
 * (AtomDump::writeDumpHeader). */
typedef struct misa_dump misa_dump;
int misa_dump_open(const char* path, misa_dump** out);
int misa_dump_frame(misa_dump* d, const misa_geom* g, const void* atoms, const void* inter, int64_t n_inter,
                    uint64_t time_step);
/* The same for a handle's resident state (downloaded first). */
int misa_dump_frame_state(misa_dump* d, misa_handle* h, uint64_t time_step);
int misa_dump_close(misa_dump* d, int64_t* records);
/* Debug mode (OutputDump::dump, output_dump.cpp:40-70): "print atoms", then
 * "id x y z" of every occupied sub-box site, "print inter", then the inter
 * atoms; numbers as std::ostream prints them by default (%g). */
int misa_dump_text(const char* path, const misa_geom* g, const void* atoms, const void* inter, int64_t n_inter);
int misa_dump_text_state(misa_handle* h, const char* path);

/* ---- frontend: YAML configuration and the stage machine ----
 *
 * ConfigValues (frontend/config_values.h:29-109) as parsed by ConfigParser
 * (frontend/config_parser.cpp:46-343) from the reference's YAML format
 * (example/config.yaml), with the same defaults and error messages; and
 * MDSimulation's stage machine (frontend/md_simulation.cpp:40-115) over the
 * resident step: stage step lengths, recoils (when an origin dump path is
 * set, as in the reference), rescales, atom dumps and thermo lines, with the
 * steps between two such events in one misa_step call. */
#define MISA_MAX_STAGES 64
#define MISA_PATH_MAX 512
#define MISA_DUMP_DEBUG 0 /* OutputMode::DEBUG: text dumps dump_<rank>_<step>.atom */
#define MISA_DUMP_COPY 1  /* OutputMode::COPY: the binary copy-mode dump */
typedef struct misa_stage {
  uint64_t steps;
  double step_length;      /* 0: the default step length */
  int32_t collision_set;   /* set_v */
  int32_t collision_lat[4];
  uint64_t collision_step; /* relative to the stage; the recoil precedes stage step collision_step - 1 */
  double pka_energy;
  double direction[3];
  int32_t rescale_set;
  int32_t pad_;
  double rescale_t;
  uint64_t rescale_every;
} misa_stage;
typedef struct misa_config {
  int64_t phase_space[3];
  double cutoff_radius_factor;
  double lattice_const;
  uint64_t time_steps; /* sum of the stages' steps */
  double time_step_length;
  int32_t create_phase;
  int32_t create_seed;
  double create_t_set;
  char read_phase_filename[MISA_PATH_MAX];
  int32_t alloy_create_seed;
  int32_t alloy_ratio[3];
  char potential_file_type[32];
  char potential_filename[MISA_PATH_MAX];
  int32_t atoms_dump_mode;
  int32_t out_by_frame;
  uint64_t atoms_dump_interval;
  char atoms_dump_file_path[MISA_PATH_MAX];
  char origin_dump_path[MISA_PATH_MAX];
  uint64_t thermo_interval;
  int32_t logs_mode; /* 0 console, 1 file */
  int32_t n_stages;
  char logs_filename[MISA_PATH_MAX];
  misa_stage stages[MISA_MAX_STAGES];
} misa_config;
typedef struct misa_md_report {
  uint64_t steps;
  double phy_time;
  int64_t stages_run, collisions, rescales, dump_frames, dump_records, thermo_lines;
  misa_thermo thermo; /* at the end */
} misa_md_report;
/* Parse a configuration file (or text); `err` receives the message. */
int misa_config_parse(const char* path, misa_config* out, char* err, int64_t err_cap);
int misa_config_parse_text(const char* text, misa_config* out, char* err, int64_t err_cap);
/* Build the world of `c` (creation mode, setfl potential) on `device`, run
 * every stage in `precision` (MISA_PREC_*), dump and log as configured. */
int misa_md_simulate(const misa_config* c, int device, int32_t precision, misa_md_report* report);

/* ---- host-only set-up: no device access, usable without a GPU ---- */

/* NeighbourIndex::make for geometry `g` (which = 0..3 as misa_get_offsets;
 * 4, 5 = pruned parity-plane slot deltas of the even / odd centres). */
int misa_host_offsets(const misa_geom* g, int which, int64_t* out, int64_t cap, int64_t* n);
/* Interpolation table (as misa_get_potential_table) straight from a file. */
int misa_host_potential_table(const char* setfl_path, int kind, int idx, double* out, int64_t cap, int64_t* n);
/* WorldBuilder::build into a caller-provided ghost-extended AtomElement
 * array of (2(n0+2g)) (n1+2g) (n2+2g) slots. */
int misa_host_build_world(const misa_geom* g, uint32_t seed, double tset, const int32_t ratio[3], uint32_t alloy_seed,
                          void* atoms);

#ifdef __cplusplus
}
#endif

#endif /* MISA_EAM_H */
