/*
 * gfmd_hip.h -- C ABI of libgfmd_hip.so, the MI355X (gfx950) elastic-force
 * hot path behind LAMMPS `fix gfmd` (Atomistica/user-gfmd).
 *
 * One handle = one GFMD surface: an nx x ny grid with ndof = 3*nu degrees of
 * freedom per grid point.  The library keeps the stiffness Phi(q), the
 * spectral work buffer, the energy partials and (for the relaxation API)
 * the solver state resident in device memory.  Calls on one handle are
 * serialised on the handle's HIP stream; distinct handles are independent.
 *
 * Field layout (host and device): a contiguous [ndof][nx][ny] block of
 * doubles, element (idof, ix, iy) at idof*nx*ny + ix*ny + iy -- the layout
 * of the reference's u_xy / f_xy buffers (fix_gfmd.cpp:613-616, y fastest).
 * A host caller holding the reference's double** passes buffer[0].
 *
 * Conventions follow GFMDSolverStatic (src/solvers/gfmd_solver_static.cpp):
 * forward transform sum_r u(r) exp(-i q.r), unnormalised inverse, Phi
 * pre-divided by nx*ny, energy summed over the full grid, Gamma-point
 * linear forces linf[i] on the z component of layer i.
 *
 * Every function returns a status code (0 = GFMD_OK); on failure
 * gfmd_last_error() returns a message for the calling thread.  The library
 * never aborts the process.
 */
#ifndef GFMD_HIP_H
#define GFMD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFMD_OK 0
#define GFMD_ERR_ARG 1         /* invalid argument */
#define GFMD_ERR_HIP 2         /* HIP runtime error */
#define GFMD_ERR_STATE 3       /* call out of order (e.g. no kernel set) */
#define GFMD_ERR_UNSUPPORTED 4 /* configuration not supported */
#define GFMD_ERR_NONFINITE 5   /* non-finite value detected */
#define GFMD_ERR_KERNEL 6      /* stiffness kernel failed (singular, ...) */
#define GFMD_ERR_CONTACT 7     /* contact model failure (atom inside sphere) */

/* Arithmetic of the bulk FFT and Phi.u work.  Energies, u0 and all
   reductions are accumulated in FP64 in both modes. */
#define GFMD_FP64 0
#define GFMD_FP32 1

typedef struct gfmd_solver gfmd_solver_t;

/* Message describing the last failure on this thread ("" if none). */
const char *gfmd_last_error(void);

/* Library version string. */
const char *gfmd_version(void);

/* Create a solver for an nx x ny grid with ndof (multiple of 3, <= 24)
   degrees of freedom per point.
   Replaces GFMDSolver::set_grid_size (src/main/gfmd_solver.cpp:79-136) and
   GFMDSolverFFT::set_grid_size (src/solvers/gfmd_solver_fft.cpp:66-93) for a
   single rank owning the whole grid. */
int gfmd_create(int nx, int ny, int ndof, int precision, gfmd_solver_t **out);

/* Release all device memory of the handle.  NULL is accepted. */
int gfmd_destroy(gfmd_solver_t *s);

/* Tabulate Phi(q) on the device from a stiffness-kernel description in the
   reference's `fix gfmd` syntax, e.g. name = "sc100", args = "height 128".
   Supported:
     sc100 [height H]   (sc100_stiffness.cpp:140-160)
     fcc100 [x] nu k1 [kappa1] [k2 [kappa2] k3 [kappa3]] [height H]
            [linf l_1 .. l_nu] [normal +-1] [mulphase m]
                        (fcc100_stiffness.cpp:59-242; nu <= 4, ndof = 3 nu)
     nonperiodic nu myu [xprd L] [yprd L]
                        (nonperiodic_stiffness.cpp:44-270: Li & Berger
                        Green's function of one grid cell, open boundaries;
                        grids up to 4096 x 4096)
     isotropic nu myu gamma [xprd L] [yprd L]
                        (isotropic_stiffness.cpp:44-86; the domain lengths
                        of its q = 0 entry default to nx, ny and must agree)
   Also installs the kernel's Gamma-point forces (the "linf" values, zero
   when absent), as StiffnessKernel::get_force_at_gamma_point provides them.
   Replaces GFMDSolverStatic::set_kernel (gfmd_solver_static.cpp:90-136) +
   fill_phi_buffer (src/main/gfmd_misc.cpp:32-102) for those kernels.
   normalize != 0 divides by nx*ny as the reference's solver does. */
int gfmd_set_kernel(gfmd_solver_t *s, const char *name, const char *args,
                    int normalize);

/* Upload a host-computed full-grid table in the reference's fill_phi_buffer
   layout: phi[(ix*ny + iy)*ndof*ndof + i*ndof + j] as interleaved complex128
   (re, im).  `normalized` != 0 means the table is already divided by nx*ny.
   Each Phi(q) must be Hermitian within 1e-6 |tr Phi(q)|, the check of
   fill_phi_buffer (gfmd_misc.cpp:53-73); otherwise GFMD_ERR_ARG.  The
   library stores the half spectrum ky <= ny/2 of (Phi(q) + conj Phi(-q))/2,
   which is the operator the reference applies when it keeps the real part
   of its complex inverse transform, so forces and energy match it also for
   tables without Phi(-q) = conj Phi(q).  This is the route for any
   StiffnessKernel the host fix still owns (set_kernel of the reference). */
int gfmd_set_phi(gfmd_solver_t *s, const double *phi, int normalized);

/* Surface cell of an `ft` kernel description (the args of
   gfmd_set_kernel(s, "ft", args)); host only, no device or handle needed.
   Writes the atoms per cell to *nat and, when the pointers are not NULL
   and *nat <= max_atoms, the cell (row-major 3 x 3, lattice vectors as
   columns, lengths scaled by the lattice constant), the atom positions
   [nat][3] (atom 0 in the top layer) and the Gamma-point force along z per
   atom (nonzero for lj/cut only).  Grid cell (ix, iy) holds atom a at
   position + ix * column 0 + iy * column 1, its displacement in degrees of
   freedom 3a .. 3a+2.  Replaces CrystalSurface::get_cell, get_positions
   (src/main/crystal_surface.h:96-140) and FTStiffnessKernel::
   get_force_at_gamma_point (src/stiffness_kernels/ft_stiffness.cpp:
   295-305), which the reference's fix queries from its kernel object. */
int gfmd_ft_cell(const char *args, int *nat, double *cell, double *positions,
                 double *linf, int max_atoms);

/* Linear force at the Gamma point per layer (nu = ndof/3 values), as
   returned by StiffnessKernel::get_force_at_gamma_point
   (src/main/surface_stiffness.cpp:427-432).  Default: zeros. */
int gfmd_set_linf(gfmd_solver_t *s, const double *linf);

/* One elastic force evaluation on host buffers u -> f, returns the energy.
   Replaces GFMDSolverStatic::post_force (gfmd_solver_static.cpp:145-249).
   u and f must not alias (the legacy GPU solver's requirement,
   src/cuda/gfmd_solver_cuda.cpp:405-407). Synchronous. */
int gfmd_post_force(gfmd_solver_t *s, const double *u, double *f,
                    double *epot);

/* Device-resident variant: u_dev/f_dev are device buffers of the handle's
   I/O type (double for GFMD_FP64, float for GFMD_FP32), epot_dev a device
   double.  Asynchronous on the handle's stream (gfmd_synchronize waits). */
int gfmd_post_force_device(gfmd_solver_t *s, const void *u_dev, void *f_dev,
                           double *epot_dev);

/* q = 0 displacement Re u(q=0) of the last evaluation (ndof values); the
   reference's GFMDSolver::get_u0 (gfmd_solver.h:88-90). Synchronous. */
int gfmd_get_u0(gfmd_solver_t *s, double *u0);

/* ----------------------------------------------------------------------
 * Atom list <-> grid glue of `fix gfmd` for atoms resident on the device
 * (FixGFMD::pre_force, fix_gfmd.cpp:659-846; post_force / grid_to_list,
 * 870-1074; the reference's device kernels for the same step,
 * src/cuda/fix_gfmd_cuda_kernel.cu:89-365).  One process owns the whole
 * grid: no ghost atoms, no reverse communication.  All arrays below are
 * device memory; the calls are asynchronous on the handle's stream and the
 * layer shift, counts and force sums stay on the device (gfmd_fix_state
 * reads them).
 * ---------------------------------------------------------------------- */
typedef struct {
  int n;              /* local atoms */
  const int *mask;    /* [n] LAMMPS group masks, or NULL: all in the group */
  int groupbit;
  int *gid;           /* [n][3] grid ix, iy and atom index in the cell
                         (atom->gid); rewritten when the layer shifts */
  const double *x;    /* [n][3] positions */
  const double *xeq;  /* [n][3] equilibrium positions (atom->xeq) */
  const double *v;    /* [n][3] velocities or NULL */
  const int *image;   /* [n] image flags (32-bit: 10 bits per dimension,
                         offset 512) or NULL: x is already unwrapped */
  const double *mass; /* [n] per-atom masses or NULL: unit masses */
} gfmd_atoms_t;

typedef struct {
  double xeqcm[3];    /* centre of mass at equilibrium (from xeq) */
  double xcm[3];      /* centre of mass at the last pre_force */
  int xshift, yshift; /* layer shift in lattice units */
  int placed;         /* atoms gathered by the last pre_force */
  int unplaced;       /* group atoms it could not place (grid id out of
                         range, or displaced by more than 64 boxes) */
  double fsum[3];     /* grid forces given to the atoms by the last
                         post_force (before fix_center) */
  int unforced;       /* group atoms that received no force */
  double masstotal;
  int natoms;         /* atoms in the group */
} gfmd_fix_state_t;

/* FixGFMD::init + get_cm_from_xeq (fix_gfmd.cpp:305-309, 542, 1895-1926):
   box[3] the periodic box lengths, invcell the inverse surface cell
   (3 x 3 row-major; rows 0 and 1 turn a centre-of-mass displacement into
   the layer shift); records the equilibrium centre of mass from a->xeq
   and clears the shift. */
int gfmd_fix_setup(gfmd_solver_t *s, const gfmd_atoms_t *a,
                   const double box[3], const double invcell[9]);
/* FixGFMD::pre_force: centre of mass (unwrapped with the image flags),
   shift = nearest integer of invcell (xcm - xeqcm), grid ids moved by the
   shift's change and wrapped, u = x - xeq with the minimum image in x and
   y onto u_dev [ndof][nx][ny], velocities onto v_grid_dev (or NULL), both
   of the handle's I/O type (double for GFMD_FP64, float for GFMD_FP32). */
int gfmd_fix_pre_force(gfmd_solver_t *s, const gfmd_atoms_t *a, void *u_dev,
                       void *v_grid_dev);
/* FixGFMD::post_force after the solver: f_atoms[i] += f_grid at atom i's
   grid point (f_grid of the handle's I/O type, f_atoms [n][3] doubles),
   the sum of those forces, and for each fix_center[d] != 0 the mean
   sum/natoms removed from every group atom's force along d. */
int gfmd_fix_post_force(gfmd_solver_t *s, const gfmd_atoms_t *a,
                        const void *f_grid_dev, double *f_atoms_dev,
                        const int fix_center[3]);
/* The glue's device state (synchronous). */
int gfmd_fix_state(gfmd_solver_t *s, gfmd_fix_state_t *out);

/* Preconditioned gradient, GFMDSolverStatic::prec_gradient
   (gfmd_solver_static.cpp:253-269 with precondition_gradient<DEF_G>,
   gfmd_misc.h:38-133): gP = F^-1[(Phi~(q) + cavg)^-1 F[g]] per wavevector,
   F the unnormalised forward transform, F^-1 the unnormalised inverse
   (real part), Phi~ the table as stored (Phi / (nx ny) when normalised).
   cavg: ndof x ndof real symmetric matrix (row-major) or NULL for zero;
   g, gP: host [ndof][nx][ny] doubles, distinct.  The reference's general
   branch (ndof > 3) writes back only components 0-2 of each product and
   leaves the others untransformed; here every component is
   preconditioned (components 0-2 agree with the reference).  The table
   is rebuilt when cavg or Phi changes.  Synchronous.  Used by the
   reference's Phi-preconditioned minimisers (src/extras/min_cg_gfmd.cpp:
   121-140, min_tr_gfmd.cpp:113-202). */
int gfmd_prec_gradient(gfmd_solver_t *s, const double *cavg, const double *g,
                       double *gp);
/* Device variant: g_dev, gp_dev of the handle's I/O type (double for
   GFMD_FP64, float for GFMD_FP32); asynchronous on the handle's stream
   (after the first call for a given cavg, which builds the table). */
int gfmd_prec_gradient_device(gfmd_solver_t *s, const double *cavg,
                              const void *g_dev, void *gp_dev);

/* Copy the device stiffness table to the host as doubles, in the device
   layout: out[(ky*ndof*ndof + e)*nx + kx] for ky in [0, ny/2], e indexing
   the Hermitian packing (ndof diagonal reals, then Re/Im of Phi_ij, i < j,
   row-major), scaled as stored (1/(nx*ny) when normalised).  The data the
   reference writes in GFMDSolverFFT::dump_stiffness
   (src/solvers/gfmd_solver_fft.cpp:294-360), returned instead of written to
   phiIJ.{real,imag}.out files. */
int gfmd_dump_stiffness(gfmd_solver_t *s, double *out);

/* Force evaluation that allocates nothing.  No counterpart in the
   reference; for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc),
   who therefore cannot create a handle.  The same kernels as
   gfmd_post_force_device, with every input and output from the caller:
   phi_dev the packed stiffness table in the handle's layout and precision
   (T[ny/2+1][ndof*ndof][nx], as gfmd_dump_stiffness returns it, T = double
   for GFMD_FP64, float for GFMD_FP32); phi_mirror 1 only if the table
   passes the x-mirror check (gfmd_phi_mirror), else 0; linf_dev ndof/3
   device doubles or NULL; u_dev, f_dev T[ndof][nx][ny], distinct;
   epot_dev a device double; u0_dev ndof device doubles or NULL; stream a
   hipStream_t (NULL: the default stream).  Asynchronous.  The scratch
   (spectrum, exchange, energy partials, twiddle tables) is a static
   device array of GFMD_NOALLOC_SCRATCH_BYTES held by each device the
   library runs on until the process ends: power-of-two grids 64..4096
   with ndof 3 or 6 up to that size (e.g. 2048 x 2048 with ndof 3 in FP64,
   4096 x 2048 in FP32); larger grids return GFMD_ERR_ARG.  Calls that
   overlap on one device race on the shared scratch: order them (one
   stream, or events).  No hipMalloc/hipFree and no memset; nothing needs
   to start at zero (every scratch element is written before it is read). */
#define GFMD_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
int gfmd_post_force_noalloc(int nx, int ny, int ndof, int precision,
                            const void *phi_dev, int phi_mirror,
                            const double *linf_dev, const void *u_dev,
                            void *f_dev, double *epot_dev, double *u0_dev,
                            void *stream);

/* The stored stiffness at `count` wavevectors (kx[q], ky[q]), 0 <= kx <
   nx, 0 <= ky <= ny/2, unpacked into full ndof x ndof complex matrices:
   out[(q*ndof + i)*ndof + j] as interleaved (re, im), scaled as stored
   (1/(nx*ny) when normalised).  What is stored is (Phi(q) + conj
   Phi(-q))/2 (see gfmd_set_phi).  Lets a host that owns a StiffnessKernel
   compare a device-tabulated table with its own get_stiffness_matrix
   (src/main/surface_stiffness.h:111-122) without copying the whole table;
   the reference has no counterpart (its table lives in host memory). */
int gfmd_stiffness_at(gfmd_solver_t *s, int count, const int *kx,
                      const int *ky, double *out);

/* Measurement hook: runs `iters` device evaluations (as
   gfmd_post_force_device) and returns the mean duration in milliseconds of
   each sweep, timed with HIP events on the handle's stream: ms[0] row r2c,
   ms[1] the column kernel (FFT + Phi.u + energy partials + inverse FFT),
   ms[2] row c2r (one of its blocks also sums the energy partials), ms[3]
   the whole evaluation.  Fast (power-of-two) path only.  No
   counterpart in the reference (its timing is MPI_Wtime around post_force,
   fix_gfmd.cpp:872-941). */
int gfmd_time_force_device(gfmd_solver_t *s, const void *u_dev, void *f_dev,
                           double *epot_dev, int iters, double *ms);

/* Measurement hook: the same evaluation launched eagerly `iters` times
   back to back, then captured once as a HIP graph of `per_graph`
   evaluations (iters a multiple of per_graph) and replayed iters /
   per_graph times; returns the mean milliseconds per evaluation of each
   (HIP events on the handle's stream).  f_dev and epot_dev end with the
   last replay's result, which equals an eager evaluation's bit for bit.
   Fast (power-of-two) path only.  No counterpart in the reference. */
int gfmd_time_force_graph(gfmd_solver_t *s, const void *u_dev, void *f_dev,
                          double *epot_dev, int iters, int per_graph,
                          double *ms_eager, double *ms_graph);

/* Autotuner: times every compiled variant of the fast path's sweeps (row
   tile shapes x column exchange/register schedules) on this handle's
   buffers for `iters` evaluations each (three passes, each combination at
   its fastest pass).  The row variant is the one with the lowest mean over
   the column variants and the column variant the one with the lowest mean
   over the row variants (their sweeps add); where some combination does
   not exist for the grid, the fastest combination.  The choice is
   remembered for the process per (nx, ny, ndof, precision) and picked up
   by handles created later with the same shape.  Returns the chosen
   variant ids and their mean time per evaluation (ms).  Generic-path
   grids have nothing to tune (returns GFMD_OK, ms_best = 0).  No
   counterpart in the reference. */
int gfmd_autotune(gfmd_solver_t *s, int iters, int *row_var, int *col_var,
                  double *ms_best);

/* Select fast-path variants explicitly (row_var < 3, col_var < 4); results
   of all variants agree to rounding.  GFMD_ERR_ARG if the variant does not
   exist for this grid. */
int gfmd_set_variant(gfmd_solver_t *s, int row_var, int col_var);

/* The handle's HIP stream (hipStream_t) and a wait on it. */
void *gfmd_stream(gfmd_solver_t *s);
int gfmd_synchronize(gfmd_solver_t *s);

/* Symmetry of the stored stiffness table, checked on the device each time
   it is written: *mirror = 1 when Phi(-qx, qy) = M Phi(qx, qy) M with M =
   -1 on every x component (sc100, fcc100 one layer, isotropic, ...), to
   1e-12 (FP64) / 1e-6 (FP32) of the trace for a device-tabulated table and
   to 4 ulp of the trace for a table uploaded with gfmd_set_phi; the column
   sweep then reads the half kx <= nx/2 of the table.  No counterpart in the reference (it reads
   the full complex table, gfmd_solver_static.cpp:203-209). */
int gfmd_phi_mirror(gfmd_solver_t *s, int *mirror);

/* Reals per wavevector the force sweep reads from the stiffness table:
   ndof*ndof (Hermitian packing), or 6 when an ndof-3 table has Phi_xy
   real and Phi_xz, Phi_yz imaginary at every wavevector (checked on the
   device when the table is written; kernels symmetric under x -> -x and
   y -> -y separately, e.g. sc100): the sweep then reads a compact copy.
   No counterpart in the reference (full complex table,
   gfmd_solver_static.cpp:203-209). */
int gfmd_phi_entries(gfmd_solver_t *s, int *entries);

/* Grid description: nx, ny, ndof, precision, and the force path in
   fast_path: 1 = the fused power-of-two FFT sweeps, 2 = the mixed-radix FFT
   (any nx, ny up to 4096, e.g. the reference's 64x37 and 63x63 grids),
   0 = direct DFTs (a side beyond 4096 that is not a supported power of
   two). */
int gfmd_get_info(gfmd_solver_t *s, int *nx, int *ny, int *ndof,
                  int *precision, int *fast_path);

/* Bytes of device memory held by the handle (GFMDSolver::memory_usage). */
size_t gfmd_memory_usage(gfmd_solver_t *s);

/* ------------------------------------------------------------------------
 * Static relaxation on the device.
 *
 * FIRE with the reference's FixFIRE semantics (src/cuda/fix_fire.cpp:
 * 117-297: half kick, power P = v.f, mixing v = (1-mix) v + mix |v| f/|f|,
 * velocity cut and mixing reset when P < 0, dt growth and mixing decay
 * after minsteps steps with P >= 0), driving gfmd_post_force_device, with
 * the contact fused into the per-step kernels.  The surface layer is layer 0
 * (components 0..2).
 * ---------------------------------------------------------------------- */

/* FIRE parameters; names and defaults of `fix fire` (fix_fire.cpp:62-77):
   minsteps 10, incfac 1.2, decfac 0.5, mix_in 0.1, mixdec 0.99,
   max_dt 0.001, xlimit 0 (off); mass 1.  adapt_dt = 0 keeps dt fixed, as the
   reference's fix does in effect (its final_integrate ends with reset_dt on
   every fix, which restores dt, fix_fire.cpp:289-305); adapt_dt = 1 lets dt
   grow to dtmax as FIRE intends. */
typedef struct {
  double dt, dtmax, incfac, decfac, mix_in, mixdec, xlimit, mass;
  int minsteps, adapt_dt;
} gfmd_fire_params_t;

#define GFMD_CONTACT_NONE 0
#define GFMD_CONTACT_LJ_SPHERE 1 /* fix contact/sphere (LJ 12-6 wall) */
#define GFMD_CONTACT_HARDWALL 2  /* rigid surface, u_z <= z_wall */
#define GFMD_CONTACT_WALL_MAP 3  /* fix wall/map: soft wall over a map */

/* A soft wall over a height map: `fix wall/map` (src/extras/
   fix_wall_map.cpp), set up once (arguments parsed, map read or taken
   from memory, low-pass filtered, shifted to z0) with the map resident in
   device memory.  Created by gfmd_wall_map_create (below). */
typedef struct gfmd_wall_map gfmd_wall_map_t;

typedef struct {
  int type;
  /* normal force on every surface site along +z (towards the indenter):
     the load per site of a load-controlled contact (q = 0 mode) */
  double pressure;
  /* site (ix, iy) rests at (x0 + a ix, y0 + a iy, z0): a, x0, y0, z0 */
  double lattice[4];
  /* LJ sphere, the arguments of `fix contact/sphere` (fix_contact_sphere.
     cpp:60-140): centre cx, cy, cz, radius, epsilon, sigma, cutoff */
  double sphere[7];
  /* hard wall: z of the rigid surface at every site, host [nx*ny], plus a
     constant offset */
  const double *wall;
  double wall_offset;
  /* surface cells with more than one atom in contact (fcc111: two atoms
     in the top layer), LJ sphere only: the cell lengths along x and y
     (0: lattice[0] for both), the number of atoms per cell that feel the
     sphere (atoms 0 .. n-1; 0 means atom 0 only, at most 4) and their
     positions relative to the site's (x0 + ix cell_x, y0 + iy cell_y, z0).
     The reference applies fix contact/sphere to every GF atom at its own
     position (fix_contact_sphere.cpp:235-259). */
  double cell_xy[2];
  int contact_atoms;
  double atom_offset[4][3];
  /* GFMD_CONTACT_WALL_MAP: the wall, acting on the contact atoms (as the
     sphere does) at their positions (x0 + ix cell_x + offset + u, ...);
     its box must be the surface's (nx cell_x by ny cell_y) */
  const gfmd_wall_map_t *wall_map;
} gfmd_contact_t;

typedef struct {
  int steps;           /* FIRE steps taken */
  int converged;       /* residual <= tol reached */
  double residual;     /* |f_total| (projected at hard-wall contacts) */
  double fext_norm;    /* |f_ext| (0 without a load) */
  double energy;       /* elastic + contact - external work */
  double contact_energy;
  double ncontact;     /* surface sites in contact */
  double seconds;      /* wall time of the call (host transfers included) */
  double device_seconds; /* the solve alone: inputs resident on the device,
                            from the first force sweep to the last check */
} gfmd_relax_result_t;

/* Fill p with the reference defaults (dt = 0.001, dtmax = 0.001). */
void gfmd_fire_defaults(gfmd_fire_params_t *p);

/* Relax u (host, [ndof][nx][ny] doubles, in/out) to
   |f_total| / |f_ext| <= tol (or |f_total| <= tol without a load),
   checking every check_every steps, at most max_steps FIRE steps.  Runs in
   the handle's precision; with GFMD_FP32 the state lives in FP32 while all
   norms and energies are FP64 (use gfmd_relax_refine for residuals below
   the FP32 floor).  Replaces `minimize`/`fix fire` + `fix gfmd` +
   `fix contact/sphere` of the reference's static tests. */
int gfmd_relax(gfmd_solver_t *s, const gfmd_contact_t *contact,
               const gfmd_fire_params_t *fire, double *u, int max_steps,
               double tol, int check_every, gfmd_relax_result_t *res);

/* Seeded self-affine rough surface (rigid counter-surface of the contact
   workloads), generated on the device: white noise (SplitMix64 of seed and
   site index, Box-Muller), filtered in Fourier space with sqrt(C(q)),
   C(q) = (q/q_l)^(-2(1+H)) for 2pi/lambda_long = q_l < q <= q_s =
   2pi/lambda_short, a plateau below q_l (lambda_long <= 0: the system
   size) and zero above q_s, then scaled to the rms slope `rms_slope`
   (central differences, lattice constant a).  out: host [nx*ny] heights,
   mean zero.  No counterpart in the reference (its maps are read from
   files, fix_wall_map.cpp:1123). */
int gfmd_rough_surface(int nx, int ny, double hurst, double rms_slope,
                       double lambda_short, double lambda_long, double a,
                       unsigned long long seed, double *out);

/* fix wall/map set-up (FixWallMap::FixWallMap, fix_wall_map.cpp:74-300;
   read_map 1079-1145, filter_map 1179-1269, displace_map 1275-1340).
   args: the fix's arguments after `wall/map`, e.g.
     "flat 2.0 -1 lj93 1.0 1.0 2.5"
     "100 1.0 0.2 3.0 -1 exp 1.0 0.5 4.0"
     "rough.map top 3.0 -1 lj93/smooth 1.0 1.0 2.0 3.0 lambda 8 pendist -1"
   with the reference's maps (flat, 100 <a0> <amplitude>, 111 <a0>
   <amplitude>, or a map file), placement (bot|top), z0, normal (1 or -1),
   potentials (lj93, lj93/smooth, lj126, exp, expf: epsilon sigma cutoff
   [outer cutoff]; dugdale: epsilon sigma) and keywords (lambda, pendist,
   lowmem, disable_interpolation, dump: accepted, no dump is written; vel
   is rejected: a moving map has no static solution).  map: NULL to read
   the file named in args, or map_nx * map_ny heights (x slow) replacing
   the file.  xprd, yprd: the periodic box the map spans.  Error messages
   are the reference's. */
int gfmd_wall_map_create(const char *args, const double *map, int map_nx,
                         int map_ny, double xprd, double yprd,
                         gfmd_wall_map_t **out);
int gfmd_wall_map_destroy(gfmd_wall_map_t *w);
/* info[6]: the map's minimum, maximum, rms height and average after the
   shift to z0, and its size nx, ny (map files and tables; 0 for the
   analytic maps) -- the fix's set-up report (fix_wall_map.cpp:1341-1350). */
int gfmd_wall_map_info(const gfmd_wall_map_t *w, double *info);
/* FixWallMap::pre_force (fix_wall_map.cpp:686-811) on n atoms on the
   device: x [n][3] host positions (z reset in place where the pendist rule
   applies), f [n][3] host forces (the wall's added), fmap[7] the fix's
   energy and vector (energy, force sums x, y, z, atoms within the cutoff,
   repulsive, attractive: compute_scalar / compute_vector, 815-845).
   GFMD_ERR_CONTACT for lj126 atoms closer than 0.01 cutoff (the
   reference's "Atom too close"). */
int gfmd_wall_map_force(gfmd_wall_map_t *w, int n, double *x, double *f,
                        double *fmap);

/* Hard-wall contact under a normal load, solved for the contact pressure
   by constrained conjugate gradients (Polonsky & Keer 1999, fixed-approach
   form: the load enters as the uniform displacement G_zz(q=0) p0 of the
   finite-height substrate).  Same problem and acceptance test as gfmd_relax
   with GFMD_CONTACT_HARDWALL: u_z <= wall + wall_offset on the surface
   layer, load `pressure` per surface site in +z, converged when the
   projected total force satisfies |f| <= tol |f_ext| (FP64).  The minimum
   of this convex problem is unique, so FIRE and this solver reach the same
   state; this one needs far fewer force-sized sweeps (two scalar
   convolutions per iteration).  Extends the reference's minimisers
   (src/extras/min_cg_gfmd.cpp: CG with the (Phi + C)^-1 preconditioner of
   gfmd_misc.h:38-133) to the bound constraint of a hard wall.
     s64    FP64 handle (required): final iterations and the residual test
     s32    FP32 handle of the same grid or NULL: iterations until its
            round-off floor, handed to s64 as the starting pressures
     u      out: [ndof][nx][ny] FP64 displacement of the final state
     p_init optional [nx][ny] starting pressures (warm start), or NULL
     iters  optional out [2]: iterations in FP32 and in FP64
   res->steps counts all iterations; the other fields as for gfmd_relax. */
int gfmd_contact_solve(gfmd_solver_t *s64, gfmd_solver_t *s32,
                       const gfmd_contact_t *contact, double *u,
                       const double *p_init, int max_iter, double tol,
                       gfmd_relax_result_t *res, int *iters);

/* An ensemble of independent hard-wall contacts (grids may differ) solved
   together: every iteration of every member is one branch of a single
   captured hipGraph (each member on its handle's stream, tiles as tuned
   for its shape), replayed until all members reach tol; FP32 stage for all
   members first when s32 is given, then FP64.  Per member as
   gfmd_contact_solve; res[i].device_seconds is the whole batch's time.
     s64[count], s32[count] (or NULL), contacts[count], u[count] (out). */
int gfmd_contact_solve_batch(int count, gfmd_solver_t *const *s64,
                             gfmd_solver_t *const *s32,
                             const gfmd_contact_t *contacts, double *const *u,
                             int max_iter, double tol,
                             gfmd_relax_result_t *res);

/* An ensemble: `members` surfaces of one grid sharing the stiffness of
   s64 (and s32, an FP32 handle of the same grid, for an FP32 stage), with
   their fields contiguous per member so that every sweep of every member
   is one launch (the member is the second grid dimension).  The handles
   must outlive the ensemble; their own contact-solver state is used for
   the per-member residual tests.  Power-of-two grids 64..4096. */
typedef struct gfmd_ensemble gfmd_ensemble_t;
int gfmd_ensemble_create(gfmd_solver_t *s64, gfmd_solver_t *s32, int members,
                         gfmd_ensemble_t **out);
int gfmd_ensemble_destroy(gfmd_ensemble_t *e);

/* Hard-wall contacts of all members of `count` ensembles (grids may differ
   between ensembles) relaxed together: one captured hipGraph per iteration
   holds one batched iteration of every ensemble; members that reach tol
   are switched off by a device flag.  contacts, u and res are flattened
   over the ensembles in order (ensemble 0's members first).  Same problem,
   stages and acceptance test per member as gfmd_contact_solve;
   res[i].device_seconds is the whole solve's time. */
int gfmd_ensemble_contact_solve(int count, gfmd_ensemble_t *const *ens,
                                const gfmd_contact_t *contacts,
                                double *const *u, int max_iter, double tol,
                                gfmd_relax_result_t *res);

/* ---------------------------------------------------------------------
   Learned warm start: a Fourier neural operator trained on contact
   solutions of this solver and run as the starting point of
   gfmd_contact_solve (no counterpart in the reference).  Architecture:
   inputs (h / h_rms, log10 p0 + 2.5) per site; lift 2 -> 32 channels; 4
   layers z = K(v) + v W + b, v = gelu(z) (no gelu after the last); K = the
   spectral convolution on the retained modes kx in [-16, 16), ky in
   [0, 16] (complex 32 x 32 weights per mode); projection 32 -> 3
   (displacement).  All dense products (pointwise mixing and the truncated
   transforms as DFT-matrix products, forward and backward) run as BF16
   MFMA matrix products with FP32 accumulation; parameters, gradients and
   Adam state are FP32.  Tensors are channels-last [batch][nx][ny][ch]; the
   flat parameter order is P (2 x 32), bP, then per layer W (32 x 32), b,
   R (32 x 17 x 32 x 32 x re/im), then Q (32 x 3), bQ.
   --------------------------------------------------------------------- */
typedef struct gfmd_fno gfmd_fno_t;
size_t gfmd_fno_param_count(void);
int gfmd_fno_create(int nx, int ny, int max_batch, gfmd_fno_t **out);
int gfmd_fno_destroy(gfmd_fno_t *f);
int gfmd_fno_set_params(gfmd_fno_t *f, const float *params);  /* host */
int gfmd_fno_get_params(gfmd_fno_t *f, float *params);        /* host */
/* in [batch][nx][ny][2] -> out [batch][nx][ny][3]; host or device
   pointers (synchronous) */
int gfmd_fno_forward(gfmd_fno_t *f, int batch, const float *in, float *out);
/* asynchronous on the network's stream (device pointers) */
int gfmd_fno_forward_device(gfmd_fno_t *f, int batch, const float *in,
                            float *out);
/* mean-squared loss against target [batch][nx][ny][3] and its gradient
   (flat, FP32) without changing the parameters */
int gfmd_fno_gradients(gfmd_fno_t *f, int batch, const float *in,
                       const float *target, double *loss, float *grad);
/* one Adam step (beta 0.9, 0.999, eps 1e-8) on a batch; loss may be NULL
   (then nothing is waited for) */
int gfmd_fno_train_step(gfmd_fno_t *f, int batch, const float *in,
                        const float *target, float lr, double *loss);
int gfmd_fno_synchronize(gfmd_fno_t *f);
/* The network's matrix-product kernel alone (device pointers; strides
   {batch, row, column} in elements): C = A B (+ C), A and B BF16 bits, C
   FP32.  For testing the kernel against a host product. */
int gfmd_fno_gemm_test(int M, int N, int K, int batch, const uint16_t *A,
                       const long long *sa, const uint16_t *B,
                       const long long *sb, float *C, const long long *sc,
                       int accumulate);

/* Damped velocity-Verlet molecular dynamics of the surface under the
   GFMD forces (plus the contact's load and LJ sphere; hard walls are not
   smooth and are rejected).  Replaces the reference's MD path: LAMMPS
   `fix nve` on the `fix gfmd` forces
   (tests/TEST_energy_conservation_fcc100/lammps.in), whose own dynamic solver is not part of the build
   (Install.sh:290-296).  Per step h = dt:
     v(n+1/2) = v(n) + h/2 (F(n)/m - damping v(n));  u(n+1) = u(n) + h v(n+1/2)
     v(n+1)   = (v(n+1/2) + h/2 F(n+1)/m) / (1 + damping h/2)
   (velocity Verlet when damping = 0).  u, v: host [ndof][nx][ny] FP64,
   in/out.  history (may be NULL): [nsteps/record_every + 1][3] FP64 of
   kinetic, potential (elastic + contact - load work) and total energy of
   the states 0, record_every, 2 record_every, ...  The step runs as one
   captured hipGraph (force sweeps + fused update + energy record). */
typedef struct {
  double dt;        /* time step */
  double mass;      /* mass per degree of freedom */
  double damping;   /* viscous damping rate (0: energy-conserving) */
  int record_every; /* energies recorded every k states (>= 1) */
} gfmd_dyn_params_t;

typedef struct {
  int steps;        /* steps integrated */
  double ekin;      /* kinetic energy of the final state */
  double epot;      /* potential energy of the final state */
  int ncontact;     /* sites inside the LJ cutoff in the final state */
  double seconds;   /* wall time of the call */
} gfmd_dyn_result_t;

int gfmd_dynamics(gfmd_solver_t *s, const gfmd_contact_t *contact,
                  const gfmd_dyn_params_t *prm, double *u, double *v,
                  int nsteps, double *history, gfmd_dyn_result_t *res);

/* Mixed-precision relaxation: FP64 outer iterations (force, residual and
   the update of u in FP64) around FP32 inner FIRE solves of the correction
   problem -- f_el(d) + r = 0 with r the FP64 residual force and, for a hard
   wall, d_z <= g the FP64 gap -- which is exact for the linear elastic
   response.  s64 (GFMD_FP64) and s32 (GFMD_FP32) hold the same grid and
   stiffness.  Stops at |f| <= tol |f_ext| (FP64) or after max_outer
   corrections; each inner solve stops when its (projected) residual has
   dropped by the factor inner_tol or is below tol |f_ext| / 2 (so the last
   correction is not solved past the goal), and the next one continues with its
   velocities and FIRE scalars (dt, mixing, cut).  res->steps counts inner
   FIRE steps; *outer_iters the corrections.  Contact: none, hard wall, or
   a nonlinear contact (LJ sphere, wall map), whose correction problem is
   f_el(d) + R + f_c(u64 + d) - f_c(u64) = 0 with R the total FP64 residual
   at u64 and both contact forces evaluated in FP64 from the FP64 state (the
   FP32 correction d only adds to it), so the FP32 rounding acts on the
   shrinking residual, not on the contact forces.  No counterpart in the
   reference (all FP64). */
int gfmd_relax_refine(gfmd_solver_t *s64, gfmd_solver_t *s32,
                      const gfmd_contact_t *contact,
                      const gfmd_fire_params_t *fire, double *u,
                      int max_outer, double tol, double inner_tol,
                      int inner_max_steps, int check_every,
                      gfmd_relax_result_t *res, int *outer_iters);

#ifdef __cplusplus
}
#endif

#endif /* GFMD_HIP_H */
