/*
 * mcfd.h — C-ABI of the MI355X (gfx950) residual hot path.
 *
 * Plain C types only: device pointers are `double*` / `int32_t*`, streams are
 * `hipStream_t` passed as `void*`. Every call returns an int status (0 = ok,
 * negative = error; `mcfd_last_error` gives the message). Calls are
 * stream-ordered and never synchronise the host, except `mcfd_status` and
 * `mcfd_sync`.
 *
 * Ownership: the caller owns every state buffer (U, R, dt). A mesh handle owns
 * the device copy of the geometry and its scratch buffers; it is used by one
 * stream at a time.
 *
 * The Python seams these entry points replace are plain module functions in
 * the reference's `src/meshcfd/core/solver.py`; a ctypes binding
 * (meshcfd_amd/native.py) calls them from the solver driver.
 */
#ifndef MCFD_H
#define MCFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCFD_ABI_VERSION 1

/* Flux-face kinds (face_kind). */
enum {
  MCFD_FACE_INTERIOR = 0,
  MCFD_FACE_PERIODIC = 1,
  MCFD_FACE_WALL = 2,          /* Wall and Symmetry: mirrored velocity */
  MCFD_FACE_VELOCITY_INLET = 3,
  MCFD_FACE_PRESSURE_OUTLET = 4,
  MCFD_FACE_FAR_FIELD = 5,
  MCFD_FACE_TRANSMISSIVE = 6,
  MCFD_FACE_PRESSURE_INLET = 7,   /* these three need mcfd_mesh_set_inflow */
  MCFD_FACE_MASS_FLOW_INLET = 8,
  MCFD_FACE_MASS_FLOW_OUTLET = 9
};

/* Accumulation modes (low 2 bits of acc_item; face index in the high bits). */
enum {
  MCFD_ACC_OWNER = 0,      /* interior face, owner side:  (-L/A) * F          */
  MCFD_ACC_NEIGHBOUR = 1,  /* interior face, neighbour:   ( L/A) * F          */
  MCFD_ACC_BOUNDARY = 2,   /* boundary / periodic a side: -(F*L)/A            */
  MCFD_ACC_PERIODIC_B = 3  /* periodic b side:             (F*L)/A            */
};

/* Error codes reported by mcfd_status (bit i of the mask = code i). */
enum {
  MCFD_ERR_STATE = 1,          /* "Euler state has non-finite or non-positive density or pressure." */
  MCFD_ERR_HLLC = 2,           /* "HLLC flux produced a non-finite value." */
  MCFD_ERR_ROE_SOUND = 3,      /* "Roe average has a non-positive sound speed." */
  MCFD_ERR_ROE = 4,            /* "Roe flux produced a non-finite value." */
  MCFD_ERR_FAR_SOUND = 5,      /* "Far-field characteristics produced invalid sound speed." */
  MCFD_ERR_FAR_STATE = 6,      /* "Far-field characteristics produced an invalid state." */
  MCFD_ERR_RESIDUAL = 7,       /* "Spatial residual contains non-finite values." */
  MCFD_ERR_LOCAL_DT = 8,       /* "Local stable time steps must be positive and finite." */
  MCFD_ERR_GLOBAL_DT = 9,      /* "The stable time step is non-positive or non-finite." */
  MCFD_ERR_TEMPERATURE = 10,   /* "Viscous flow requires finite positive temperature." */
  MCFD_ERR_VISCOSITY = 11,     /* "Sutherland law produced an invalid viscosity." */
  MCFD_ERR_INLET_OUTFLOW = 12, /* "Pressure inlet cannot prescribe a supersonic outflow." */
  MCFD_ERR_INLET_SUPERSONIC = 13, /* "Pressure inlet needs Mach number to define a supersonic inflow." */
  MCFD_ERR_INLET_TOTALS = 14,  /* "Pressure inlet total conditions are incompatible with the interior characteristic state." */
  MCFD_ERR_INLET_STATE = 15,   /* "Pressure inlet produced an invalid state." */
  MCFD_ERR_MASS_PRESSURE = 16, /* "Mass-flow inlet has invalid interior pressure." */
  MCFD_ERR_MASS_FLOW = 17,     /* "Mass-flow inlet cannot satisfy the configured flow rate." */
  MCFD_ERR_OUTLET_DENSITY = 18 /* "Mass-flow outlet has invalid interior density." */
};

/* Molecular viscosity models (McfdSettings.viscosity), solver.py:2721-2737. */
enum {
  MCFD_VISCOSITY_INVISCID = 0,
  MCFD_VISCOSITY_CONSTANT = 1,
  MCFD_VISCOSITY_SUTHERLAND = 2
};

/*
 * Geometry in device layout (host pointers; copied by mcfd_mesh_create).
 * All per-cell lists are CSR (ptr has n_cells+1 entries) and hold their
 * entries in the reference's summation order.
 */
typedef struct McfdMeshDesc {
  int32_t n_cells;
  int32_t n_faces;          /* flux faces: interior, periodic pairs, boundary */
  int32_t n_grad;           /* LSQ directed edges */
  int32_t n_lim;            /* limiter entries */
  int32_t n_acc;            /* accumulation items */
  int32_t n_dts;            /* spectral-radius items */
  int32_t n_bc;             /* boundary records */
  int32_t reserved;
  const double* area;       /* [n_cells] */
  const int32_t* grad_ptr;  /* [n_cells+1] */
  const int32_t* grad_nb;   /* [n_grad] neighbour cell */
  const double* grad_cx;    /* [n_grad] LSQ coefficient, x */
  const double* grad_cy;    /* [n_grad] LSQ coefficient, y */
  const int32_t* lim_ptr;   /* [n_cells+1] */
  const double* lim_dx;     /* [n_lim] face midpoint - centroid */
  const double* lim_dy;
  const int32_t* face_left;   /* [n_faces] owner (periodic: owner of face a) */
  const int32_t* face_right;  /* [n_faces] neighbour (periodic: owner of b), -1 on boundary */
  const int32_t* face_kind;   /* [n_faces] MCFD_FACE_* */
  const int32_t* face_bc;     /* [n_faces] boundary record or -1 */
  const double* face_nx;      /* [n_faces] unit normal (periodic: of face a) */
  const double* face_ny;
  const double* face_len;     /* [n_faces] length */
  const double* face_lx;      /* [n_faces] left offset (face midpoint - left centroid) */
  const double* face_ly;
  const double* face_rx;      /* [n_faces] right offset */
  const double* face_ry;
  const int32_t* acc_ptr;     /* [n_cells+1] */
  const int32_t* acc_item;    /* [n_acc] face*4 + MCFD_ACC_* */
  const int32_t* dts_ptr;     /* [n_cells+1] */
  const double* dts_nx;       /* [n_dts] face normal */
  const double* dts_ny;
  const double* dts_len;      /* [n_dts] face length */
  const int32_t* dts_bc;      /* [n_dts] boundary record or -1 */
  const int32_t* bc_kind;     /* [n_bc] MCFD_FACE_* */
  const double* bc_state;     /* [n_bc*4] exterior primitive; outlet uses [3] = p */
  const double* bc_wall;      /* [n_bc*2] wall velocity */
  const double* acc_coef;     /* [n_acc] -L/A (owner) or L/A (neighbour) of interior
                                 items, the reference's CSR data values; 0 otherwise */
} McfdMeshDesc;

/*
 * Cell blocks for the fused stage kernel (host pointers; copied). Block b owns
 * the consecutive cells [own_start, own_start + n_own) (the blocks tile the
 * mesh in order, at most block_cells cells each), has at most 256 own +
 * ring-1 cells and touches at most 256 flux faces; its local cells are own
 * cells, then ring-1 and ring-2 halo cells. Everything a block reads besides
 * cell states is packed block by block, so each phase of the kernel reads it
 * with one coalesced load per item:
 *   info[b*12] = {cells_start, n_own, n_ring1, n_ring2, faces_start, n_faces,
 *                 gnb_start, gptr_start, own_start, lent_start, 0, 0};
 *   gptr / lptr: per block n_own + n_ring1 + 1 block-relative offsets of each
 *     own / ring-1 cell's LSQ entries (gnb, gent) and limiter entries (lent);
 *   frec: one 64-byte record per block face,
 *     {nx, ny, lx, ly, rx, ry, length, kind | bc << 4, left | right << 16}
 *     (local left/right cell ids, 0xFFFF: none; bc -1: none).
 */
typedef struct McfdBlockDesc {
  int32_t n_blocks;
  int32_t block_cells;
  int32_t max_local;        /* max local cells of a block */
  int32_t max_grad;         /* max own + ring-1 cells of a block */
  int32_t max_faces;        /* max faces of a block */
  int32_t n_local;          /* length of cells */
  int32_t n_gptr;           /* length of gptr and lptr */
  int32_t n_gnb;            /* length of gnb (gent has twice as many doubles) */
  int32_t n_lent;           /* limiter entries (lent has twice as many doubles) */
  int32_t n_block_faces;    /* length of faces and frec */
  const int32_t* info;      /* [n_blocks*12] */
  const int32_t* cells;     /* local -> global cell */
  const int32_t* gptr;
  const uint16_t* gnb;      /* LSQ neighbours as local cell ids, reference entry order */
  const double* gent;       /* [n_gnb*2] LSQ coefficients (cx, cy) */
  const int32_t* lptr;
  const double* lent;       /* [n_lent*2] limiter offsets (dx, dy), reference entry order */
  const int32_t* faces;     /* local face -> flux face */
  const void* frec;         /* [n_block_faces] 64-byte face records */
  const uint16_t* acc_local;/* [n_acc] accumulation items as local face slot*4 + mode */
  const uint16_t* dts_local;/* [n_dts] spectral-radius items as local face slot, | 0x8000 when
                               the item's normal/length/exterior state is read from the
                               global item (periodic b sides, inlet/outlet/far-field faces) */
} McfdBlockDesc;

typedef struct McfdSettings {
  int32_t flux;            /* 0 Rusanov, 1 HLLC, 2 Roe */
  int32_t reconstruction;  /* 0 first order, 1 MUSCL linear, 2 MUSCL parabolic */
  int32_t limiter;         /* 0 Barth-Jespersen, 1 Minmod, 2 Van Leer, 3 Venkatakrishnan */
  int32_t viscosity;       /* MCFD_VISCOSITY_*; non-zero needs mcfd_mesh_set_viscous */
  double gamma;
  double cfl;
  double acceleration_x;
  double acceleration_y;
  /* laminar viscous terms (NumericalSettings fields of the same names) */
  double gas_constant;
  double prandtl_number;
  double dynamic_viscosity;               /* Constant model */
  double sutherland_reference_temperature;
  double sutherland_reference_viscosity;
  double sutherland_temperature;
} McfdSettings;

/* Stage update ops (McfdStage.op). h = dt (scalar dt[0], or dt[i] if dt_per_cell). */
enum {
  MCFD_OP_EULER = 0,   /* out = U + (c0*h)*R                                   */
  MCFD_OP_SSP = 1,     /* out = c0*U0 + c1*(U + h*R)                           */
  MCFD_OP_RK4 = 2,     /* out = c0*U0 + c1*U + (c2*h)*R                        */
  MCFD_OP_RK4_FINAL = 3, /* out = c0*U0 + c1*U1 + (c2*h)*R1 + c3*U + (c4*h)*R  */
  MCFD_OP_RESIDUAL = 4 /* R only (written to r_out)                            */
};

typedef struct McfdStage {
  int32_t op;
  int32_t dt_per_cell;
  const double* u;       /* state the residual is evaluated at, (C,4) */
  const double* u0;      /* (C,4) or NULL */
  const double* u1;      /* (C,4) or NULL (RK4 final: stage three) */
  const double* r1;      /* (C,4) or NULL (RK4 final: residual at stage three) */
  const double* dt;      /* device: [1] or [C] */
  double* out;           /* (C,4) updated state, may be NULL for MCFD_OP_RESIDUAL */
  double* r_out;         /* (C,4) residual copy, or NULL */
  double c[5];
  /* Fused path only: with dt_min set, also compute the next step's local
   * stable steps from the updated state, their minimum into dt_min[0]
   * (device) and, when dt_next is set, the steps themselves into
   * dt_next [C]. Or both NULL. */
  double* dt_next;
  double* dt_min;
} McfdStage;

/*
 * Implicit-Euler preconditioner structure (host pointers; device cell ids).
 * Records are face pairs (interior faces, periodic pairs) and boundary faces
 * (every face with neighbour < 0, periodic ones included, as the reference
 * counts them again in the diagonal). Per cell: diagonal items in the
 * reference's add.at order; off-diagonal items sorted by the natural
 * (mesh-order) index of their column, lower columns first, off_split[i] the
 * first upper item. Item = record*2 + side (0 owner side, 1 neighbour side).
 */
typedef struct McfdImplicitDesc {
  int32_t n_records;
  int32_t launch_path;         /* LU-SGS sweeps: 0 chosen per mesh (by level sizes, or by timing
                                  both paths on the first application, which synchronises the
                                  stream), 1 one workgroup, 2 one launch per level; every path
                                  gives the same bits */
  const int32_t* rec_o;
  const int32_t* rec_n;        /* -1 for boundary records */
  const double* rec_nox;       /* normal paired with the owner-side state */
  const double* rec_noy;
  const double* rec_nnx;       /* normal paired with the neighbour-side state */
  const double* rec_nny;
  const double* rec_len;
  const int32_t* diag_ptr;     /* [n_cells+1] */
  const int32_t* diag_item;    /* record*2 + side; -(record*2 + side) - 1 for the entries of a
                                  periodic pair whose two faces share one owner, which the
                                  reference's sparse assembly sums onto the diagonal */
  const int32_t* off_ptr;      /* [n_cells+1] */
  const int32_t* off_split;    /* [n_cells] */
  const int32_t* off_item;
  const int32_t* off_nb;       /* device cell of the column */
  const int32_t* natural_order;/* [n_cells] device cell at each natural (mesh) position */
  const int32_t* off_t;        /* per off-diagonal item: the column cell's item for the same
                                  record (its coefficient is the transposed entry) */
} McfdImplicitDesc;

/* Matrix-free implicit operator and GMRES arguments (device pointers). */
typedef struct McfdGmres {
  const double* state;    /* (C,4) Newton iterate s */
  const double* rate;     /* (C,4) first-order residual at s */
  const double* rhs;      /* (C,4) right-hand side b */
  double* x;              /* (C,4) solution */
  const double* dt;       /* [1] or [C] */
  int32_t dt_per_cell;
  int32_t krylov;         /* restart length m */
  int32_t max_iterations;
  int32_t iterations;     /* out: inner iterations used */
  int32_t info;           /* out: 0 converged, 1 not converged */
  int32_t reserved;
  double sigma[4];        /* variable scales */
  double state_norm;      /* ||state / sigma||_2; negative: computed on the device and stored here */
  double tolerance;       /* relative tolerance rtol */
  double residual;        /* out: final true residual ||b - A x|| */
} McfdGmres;

typedef struct McfdMesh McfdMesh;

/* Library / ABI version. */
int mcfd_version(void);

/* Message of the last failing call on this thread. */
int mcfd_last_error(char* buffer, size_t size);

/* Upload a mesh (replaces `_build_geometry`'s output `_Geometry`, solver.py:122-176). */
int mcfd_mesh_create(const McfdMeshDesc* desc, McfdMesh** mesh);
int mcfd_mesh_destroy(McfdMesh* mesh);

/* Upload cell blocks and switch mcfd_stage / mcfd_residual to the fused
 * one-kernel-per-stage path; NULL switches back to the four-kernel path.
 * No counterpart in the reference (its residual is whole-array NumPy passes). */
int mcfd_mesh_set_blocks(McfdMesh* mesh, const McfdBlockDesc* blocks);

/* First-order viscous stages on the fused path: per block LSQ entry (the
 * order of McfdBlockDesc.gent, n_vent entries) the viscous WLSQ weights
 * (wx, wy) as vent[2e], vent[2e+1]; per item of the viscous accumulation list
 * of mcfd_mesh_set_viscous (n_vacc items) its face as the local face slot*4
 * + mode of the block owning the item's cell. Needs mcfd_mesh_set_blocks and
 * mcfd_mesh_set_viscous first (again after either is replaced). Then a
 * first-order viscous residual or stage (`_residual` with viscous terms,
 * solver.py:2437-2718, and the implicit path's first-order Jacobian
 * products, :1532-1567) runs as one launch with the unfused kernels' bits. */
int mcfd_mesh_set_block_viscous(McfdMesh* mesh, const double* vent, int64_t n_vent, const uint16_t* vacc_local,
                                int64_t n_vacc);

/* R = residual(U) (replaces `_residual`, solver.py:2437-2718). */
int mcfd_residual(McfdMesh* mesh, const McfdSettings* settings, const double* u, double* r,
                  void* stream);

/* Per-cell stable steps and their minimum (replaces `_stable_time_steps` /
 * `_stable_time_step`, solver.py:4117-4343). dt_cell [C], dt_min [1], device. */
int mcfd_time_steps(McfdMesh* mesh, const McfdSettings* settings, const double* u,
                    double* dt_cell, double* dt_min, void* stream);

/* mcfd_time_steps from primitive states w (C,4) = (rho, u, v, p) on the device,
 * as `_stable_time_steps` takes them (solver.py:4132). */
int mcfd_time_steps_primitive(McfdMesh* mesh, const McfdSettings* settings, const double* w,
                              double* dt_cell, double* dt_min, void* stream);

/* The end-of-step state check of `solve_case` (solver.py:806-808:
 * `_primitive_from_conservative` of the new state): sets MCFD_ERR_STATE with
 * the first bad cell in the device status. */
int mcfd_check_state(McfdMesh* mesh, const McfdSettings* settings, const double* u, void* stream);

/* clock = [time, dt]: dt = min(dt_source[0], end_time - time) (dt_source NULL:
 * fixed_dt), checked positive/finite, then time += dt (solver.py:452-465, 812).
 * reset != 0 sets dt_source[0] = +inf after reading it (for dt_min of the
 * next fused stage). */
int mcfd_advance_clock(McfdMesh* mesh, const double* dt_source, double fixed_dt,
                       double end_time, double* clock, int32_t reset, void* stream);

/* One residual evaluation fused with an explicit stage update (the inlined
 * stages of `solve_case`, solver.py:532-756, and `_explicit_rk4_step`,
 * solver.py:1128-1182). */
int mcfd_stage(McfdMesh* mesh, const McfdSettings* settings, const McfdStage* stage,
               void* stream);

/* Upload the LU-SGS structure (builds the level schedules); NULL frees it. */
int mcfd_mesh_set_implicit(McfdMesh* mesh, const McfdImplicitDesc* desc);

/* Assemble S from the first-order spectral radii at state u with steps dt
 * (replaces `_implicit_preconditioner`'s assembly, solver.py:1185-1301). */
int mcfd_lusgs_assemble(McfdMesh* mesh, const McfdSettings* settings, const double* u,
                        const double* dt, int32_t dt_per_cell, void* stream);

/* out = U^-1 (D (L^-1 v)) (replaces `apply_lu_sgs`, solver.py:1315-1319).
 * On a mesh whose levels average 384-4096 cells, the first call with
 * v != out (here or through mcfd_lusgs_apply_transpose / mcfd_gmres) times
 * the one-workgroup sweeps against one launch per level, waits for the
 * stream, and keeps the faster; both give identical results. Make that
 * first call outside HIP graph capture. */
int mcfd_lusgs_apply(McfdMesh* mesh, const double* v, double* out, void* stream);

/* out = v - dt * J v / sigma, one-sided difference of the first-order residual
 * (replaces `_matrix_free_implicit_operator`'s matvec, solver.py:1543-1563). */
int mcfd_jacobian_apply(McfdMesh* mesh, const McfdSettings* jacobian_settings, McfdGmres* op,
                        const double* v, double* out, void* stream);

/* Restarted left-preconditioned GMRES on the matrix-free operator with the
 * assembled LU-SGS (replaces `_solve_gmres` + scipy gmres, solver.py:1489-1529).
 * Synchronises the stream once per inner iteration. */
int mcfd_gmres(McfdMesh* mesh, const McfdSettings* jacobian_settings, McfdGmres* args,
               void* stream);

/* scaled = (u - u0 - dt*r)/sigma; norms[0] = max|scaled|, norms[1] =
 * max|(u - u0)/sigma| (solver.py:1415-1424). scaled may be NULL. */
int mcfd_newton_equation(McfdMesh* mesh, const double* u, const double* u0, const double* r,
                         const double* dt, int32_t dt_per_cell, const double* sigma_host,
                         double* scaled, double* norms, void* stream);

/* out = u + (damping*corr)*sigma (solver.py:1465-1468). */
int mcfd_newton_candidate(McfdMesh* mesh, const double* u, const double* corr, double damping,
                          const double* sigma_host, double* out, void* stream);

/* out[0] = ||a||_2 over the (C,4) array (device), deterministic. */
int mcfd_norm2(McfdMesh* mesh, const double* a, double* out, void* stream);

/* MUSCL parabolic data (host pointers, device cell ids): per-cell CSR of the
 * two-ring stencil (ptr [C+1], nb [n_entries]), per entry the 5 fit
 * coefficients P_e with (grad, curvature) = sum_e P_e (W_e - W_i), and the
 * per-cell full-rank flag [C]. Replaces `_parabolic_stencil` and the per-call
 * normal-equation solve of `_parabolic_reconstruction` (solver.py:2241-2372).
 * NULL ptr frees. */
int mcfd_mesh_set_parabolic(McfdMesh* mesh, int32_t n_entries, const int32_t* ptr,
                            const int32_t* nb, const double* coef, const int32_t* full);

/*
 * Laminar viscous terms (host pointers, device cell and flux-face ids;
 * copied). Replaces `_viscous_gradients`, `_viscous_residual` and the
 * diffusive part of `_stable_time_steps` (solver.py:3615-3714, 3750-4114,
 * 4246-4332) for the Constant and Sutherland models:
 *
 * - weighted least squares of (u, v, T) per cell over its LSQ entries
 *   (grad_ptr order, weights wl_wx/wl_wy = w*dx, w*dy with
 *   w = 1/max(dx^2+dy^2, 1e-300)) followed by one ghost entry per
 *   non-periodic boundary face it owns (gh_* lists, the reference's
 *   diffusive-boundary order); wl_cell = {xx, xy, yy, 1/det} per cell,
 *   1/det = 0 where the reference's determinant test leaves the gradient 0;
 * - per boundary record: bv_code (1 wall, 2 symmetry, 3 velocity inlet,
 *   5 far field, 0 zero normal gradient), bv_flags (bit 0 no-slip, bits 1-2
 *   thermal: 1 temperature, 2 heat flux), bv_geom = {dn, tx, ty, value} with
 *   dn = (face centre - owner centroid).n and (tx, ty) = offset - dn*n;
 * - viscous accumulation per cell: items flux_face*4 + mode, mode 0 adds
 *   va_coef*F to the interior sum S (the divergence operator's CSR data),
 *   mode 1 adds (F*L)/A, mode 2 adds (-(F*L))/A to 0 - S;
 * - diffusive time-step items per cell: partner cell (-1 on boundary faces),
 *   face length and max(distance, 1e-14), summed in the reference's order.
 */
typedef struct McfdViscousDesc {
  int32_t n_ghost;
  int32_t n_vacc;
  int32_t n_vdt;
  int32_t reserved;
  const double* wl_wx;       /* [n_grad] */
  const double* wl_wy;       /* [n_grad] */
  const double* wl_cell;     /* [n_cells*4] */
  const int32_t* gh_ptr;     /* [n_cells+1] */
  const int32_t* gh_face;    /* [n_ghost] flux face */
  const double* gh_wx;       /* [n_ghost] */
  const double* gh_wy;       /* [n_ghost] */
  const int32_t* bv_code;    /* [n_bc] */
  const int32_t* bv_flags;   /* [n_bc] */
  const double* bv_geom;     /* [n_bc*4] */
  const int32_t* va_ptr;     /* [n_cells+1] */
  const int32_t* va_item;    /* [n_vacc] */
  const double* va_coef;     /* [n_vacc] */
  const int32_t* vd_ptr;     /* [n_cells+1] */
  const int32_t* vd_nb;      /* [n_vdt] */
  const double* vd_len;      /* [n_vdt] */
  const double* vd_dist;     /* [n_vdt] */
} McfdViscousDesc;

/* Upload the viscous lists and allocate their scratch (gradients, face
 * fluxes, diffusivities); NULL frees. Afterwards residual, stage and
 * time-step calls whose settings have viscosity != 0 add the viscous terms
 * (stages then run the unfused kernels). */
int mcfd_mesh_set_viscous(McfdMesh* mesh, const McfdViscousDesc* desc);

/*
 * Pressure-inlet, mass-flow-inlet and mass-flow-outlet records (host
 * pointers; copied). Replaces `_pressure_inlet_boundary_states`,
 * `_mass_flow_inlet_boundary_states` and `_mass_flow_outlet_boundary_states`
 * (solver.py:4412-4577) at their residual and time-step call sites
 * (solver.py:2626-2705, 4155-4193). The records are the n_records boundary
 * records from first_record on (kinds 7, 8, 9); per record:
 *
 * - face: its flux face;
 * - params: pressure inlet {total pressure, total temperature, dx, dy},
 *   mass-flow inlet {mass flow, total temperature, dx, dy} (d the unit flow
 *   direction), mass-flow outlet {mass flow, summed face length of the
 *   patch, 0, 0};
 * - projection: n.d (pressure inlet), -(n.d) (mass-flow inlet), 0 otherwise,
 *   as the reference evaluates them (normals @ direction);
 *
 * and per mass-flow-inlet patch p (records patch_rec[p] .. patch_rec[p+1]-1,
 * offsets from first_record): the leaves {start, size} (offsets in the
 * patch) leaf[2*(leaf_ptr[p] .. leaf_ptr[p+1]-1)] and the token program
 * program[prog_ptr[p] .. prog_ptr[p+1]-1] (t >= 0: push the sum of leaf t of
 * the patch; -1: pop b, pop a, push a + b) that evaluate the patch's flow
 * sum in NumPy's pairwise order (np.sum over a contiguous array: blocks of
 * 8192 added in turn, each split in halves rounded down to a multiple of 8
 * down to leaves of at most 128 values summed with 8 accumulators). A patch
 * may hold at most 4096 records.
 */
typedef struct McfdInflowDesc {
  int32_t first_record;
  int32_t n_records;
  int32_t n_patches;
  int32_t n_leaves;
  int32_t n_program;
  int32_t reserved;
  const int32_t* face;        /* [n_records] */
  const double* params;       /* [n_records*4] */
  const double* projection;   /* [n_records] */
  const int32_t* patch_rec;   /* [n_patches+1] */
  const int32_t* leaf_ptr;    /* [n_patches+1] */
  const int32_t* leaf;        /* [n_leaves*2] */
  const int32_t* prog_ptr;    /* [n_patches+1] */
  const int32_t* program;     /* [n_program] */
} McfdInflowDesc;

/* Upload the inflow records; NULL frees. A mesh with records of kinds 7-9
 * needs this before residual, stage and time-step calls, which then run the
 * unfused kernels (inviscid only; no adjoint). */
int mcfd_mesh_set_inflow(McfdMesh* mesh, const McfdInflowDesc* desc);

/* Upload the reverse-entry map of the LSQ lists (grad_rev[e] = entry of the
 * neighbour that points back) and allocate the adjoint scratch; NULL frees. */
int mcfd_mesh_set_adjoint(McfdMesh* mesh, const int32_t* grad_rev);

/* MUSCL parabolic adjoint: per two-ring stencil entry e (i -> j at offset
 * d) the entry of j's list that points back at i (offset -d); the stencils
 * are symmetric. Needs mcfd_mesh_set_parabolic and mcfd_mesh_set_adjoint;
 * afterwards mcfd_residual_vjp covers MUSCL parabolic settings. */
int mcfd_mesh_set_parabolic_adjoint(McfdMesh* mesh, const int32_t* par_rev);

/* ubar = lambda^T dR/dU at state u: the hand-written reverse of `_residual`
 * (solver.py:2437-2718) through fluxes, ghost states, reconstruction,
 * limiter, LSQ gradient and primitive conversion. Pressure-inlet and
 * mass-flow records (solver.py:4412-4577) are differentiated through their
 * exterior states: the pressure inlet's Mach number by the implicit-function
 * derivative of its invariant match, the mass-flow inlet's patch speed by
 * that of its flow balance (a per-patch reduction), the mass-flow outlet in
 * closed form; needs mcfd_mesh_set_inflow like the residual. If sbar is not
 * NULL it receives per boundary record [n_bc*4] lambda^T dR/d(exterior
 * state) for velocity-inlet and far-field records (0 for the others), for
 * gradients with respect to inflow conditions. No counterpart in the
 * reference (which has no adjoint). */
int mcfd_residual_vjp(McfdMesh* mesh, const McfdSettings* settings, const double* u,
                      const double* lambda, double* ubar, double* sbar, void* stream);

/* Geometry adjoint lambda^T dR/d(geometry) at state u: per flux face
 * face_bar[7k..7k+6] = (nx, ny, L, lx, ly, rx, ry)-bar (normal, length,
 * reconstruction offsets of the left and right side), per LSQ entry
 * grad_bar[2e..2e+1] = (cx, cy)-bar, per limiter entry lim_bar[2e..2e+1] =
 * (dx, dy)-bar (device order). The cell-area adjoint is -lambda.(R - body
 * force)/A and is left to the caller. Chained through the geometry's
 * dependence on the mesh nodes this gives gradients with respect to shape
 * parameters. Inviscid Rusanov or HLLC, first order or MUSCL linear, no
 * pressure-inlet or mass-flow boundaries; needs mcfd_mesh_set_adjoint. No
 * counterpart in the reference. */
int mcfd_residual_geometry_vjp(McfdMesh* mesh, const McfdSettings* settings, const double* u,
                               const double* lambda, double* face_bar, double* grad_bar,
                               double* lim_bar, void* stream);

/* Transposed LU-SGS: out = M^T v with M = U^-1 D L^-1 as assembled by
 * mcfd_lusgs_assemble (forward sweep on U^T, backward on L^T). The
 * preconditioner of the adjoint GMRES; no counterpart in the reference. */
int mcfd_lusgs_apply_transpose(McfdMesh* mesh, const double* v, double* out, void* stream);

/* Adjoint pseudo-time operator out = A^T v = v - J^T (dt v) at g->state,
 * with J the exact Jacobian of the residual for `settings` (through
 * mcfd_residual_vjp) and dt from g->dt / g->dt_per_cell. No counterpart in the
 * reference. */
int mcfd_adjoint_apply(McfdMesh* mesh, const McfdSettings* settings, McfdGmres* g, const double* v,
                       double* out, void* stream);

/* Adjoint GMRES: solves A^T x = g->rhs (A^T as in mcfd_adjoint_apply) with
 * the algorithm of mcfd_gmres, left-preconditioned by the transposed
 * LU-SGS of the matrix last assembled by mcfd_lusgs_assemble. g->rate,
 * sigma and state_norm are unused. Needs mcfd_mesh_set_implicit and
 * mcfd_mesh_set_adjoint. No counterpart in the reference. */
int mcfd_adjoint_gmres(McfdMesh* mesh, const McfdSettings* settings, McfdGmres* g, void* stream);

/* Batched cases (cells of case c in [case_ptr[c], case_ptr[c+1]), device):
 * per-cell stable steps, then per case dt = min(case minimum, end_time -
 * time_c), time_c += dt, and dt written to every cell of the case so the
 * stage kernels run with per-cell steps. clock = [time, dt] per case. No
 * counterpart in the reference (one case per run). */
int mcfd_batch_time_steps(McfdMesh* mesh, const McfdSettings* settings, const double* u,
                          int32_t n_cases, const int32_t* case_ptr, double end_time,
                          double* clock, double* dt_cell, void* stream);

/* The per-case part of mcfd_batch_time_steps alone, when dt_cell already
 * holds the per-cell stable steps (written by the final fused stage of the
 * previous step through McfdStage.dt_next): per case dt = min(case minimum,
 * end_time - time_c), time_c += dt, dt written to every cell of the case.
 * No counterpart in the reference (one case per run). */
int mcfd_batch_case_steps(McfdMesh* mesh, int32_t n_cases, const int32_t* case_ptr, double end_time,
                          double* clock, double* dt_cell, void* stream);

/* Per-component sum of R^2 over cells into sums[4] (device), deterministic. */
int mcfd_sum_squares(McfdMesh* mesh, const double* r, double* sums, void* stream);

/* mcfd_sum_squares without a mesh and without allocating: r (C,4) and
 * sums[4] on the device, n_cells <= 16,777,216 (else an error). No
 * counterpart in the reference; for callers that cannot allocate device
 * memory (HIP graph capture, serving loops). The block partials live in one
 * static 2 MiB device array per GPU, held until the process ends: calls that
 * overlap on one GPU (concurrent streams or threads) race on it. Same bits as
 * mcfd_sum_squares on a mesh of n_cells cells. */
int mcfd_sum_squares_noalloc(const double* r, int32_t n_cells, double* sums, void* stream);

/* dst[0:count] = src[0:count] (device, count even): a plain streaming copy,
 * the reference kernel for HBM byte counters (FETCH_SIZE / WRITE_SIZE). */
int mcfd_copy(const double* src, double* dst, size_t count, void* stream);

/* Synchronise `stream`, read and clear the device status: mask of MCFD_ERR_*
 * bits and the lowest failing cell/face index. Returns 0. */
int mcfd_status(McfdMesh* mesh, void* stream, int32_t* mask, int32_t* index);

/* Kernel kinds timed by mcfd_timing. */
enum {
  MCFD_KERNEL_PRIMITIVE = 0,
  MCFD_KERNEL_GRADIENT = 1,
  MCFD_KERNEL_FLUX = 2,
  MCFD_KERNEL_UPDATE = 3,
  MCFD_KERNEL_TIME_STEP = 4,
  MCFD_KERNEL_REDUCE = 5,
  MCFD_KERNEL_STAGE = 6,
  MCFD_KERNEL_VISCOUS = 7,
  MCFD_KERNEL_KINDS = 8
};

/* Per-launch HIP-event timing on the launch stream, for benchmarking.
 * enable = 1: start recording (clears totals); enable = 0: stop.
 * On return (after synchronising the recorded events) total_ms[k] and
 * counts[k], k < MCFD_KERNEL_KINDS, hold the summed durations and launch
 * counts per kernel kind since recording started. Either array may be NULL. */
int mcfd_timing(McfdMesh* mesh, int32_t enable, double* total_ms, int32_t* counts);

#ifdef __cplusplus
}
#endif

#endif /* MCFD_H */
