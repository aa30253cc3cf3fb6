// Hard-wall contact by constrained conjugate gradients (contact_cg.hip).
#pragma once

#include "gfmd_internal.h"

namespace gfmd {

// Scalars of the iteration kept in device memory.
enum {
  PK_GG = 0,  // sum of g^2 over the contact set (this iteration)
  PK_GOLD,    // the same of the previous iteration
  PK_NC,      // sites in the contact set
  PK_BETA,    // conjugacy factor of the direction update
  PK_TAU,     // step length
  PK_CONJ,    // 1 if the last update had no overlap (keep conjugacy)
  PK_DP,      // sum |dp| of the last update
  PK_PSUM,    // sum p after the last update
  PK_OVL,     // overlap sites of the last update
  PK_ITER,    // iterations done
  PK_EMPTY,   // 1 when the contact set is empty: the direction and step
              // then use the overlap set {g < 0} (projected steepest
              // descent), so that a start without pressure can progress
  PK_COUNT
};

// Fields of nx*ny values of the compute type T.
struct PKBufs {
  void *p, *t, *g, *r, *u, *h;
  double *state, *part;  // [PK_COUNT], [8][nblk]
  int nblk;
  size_t n;
  const int *gate;  // device flags [members]; 0 makes a member's
                    // iteration kernels no-ops
  int members = 1;  // ensemble: fields [members][n], state
                    // [members][PK_COUNT], part [members][8][nblk]
};

enum { PK_GAP = 0, PK_CURV_UPDATE = 1 };

// PK_GAP (after u = G*p): g = u + h, beta, t = g + beta t on {p > 0}.
// PK_CURV_UPDATE (after r = G*t): tau, projected update of p, measures.
void launch_pk(const PKBufs &b, bool fp64, int which, hipStream_t st);
// p = p_init (T, clamped at 0) or the uniform p0; t = 0; state cleared;
// the gate flag set to 1.
// p0m (device, [members]) overrides p0 per member.
void launch_pk_init(const PKBufs &b, bool fp64, const void *p_init, double p0,
                    hipStream_t st, const double *p0m = nullptr);
// h = wall + offset - c (wall: T[members][nx*ny]); cm (device, [members])
// overrides c per member.
void launch_pk_h(const PKBufs &b, bool fp64, const void *wall, double offset,
                 double c, hipStream_t st, const double *cm = nullptr);
// tr = T[d][nx*ny]: component zc = p0 - p, others 0
void launch_pk_traction(const PKBufs &b, bool fp64, int d, int zc, double p0,
                        void *tr, hipStream_t st);
// u_zc = wall + offset where p > 0 or where u_zc exceeds it
void launch_pk_snap(const PKBufs &b, bool fp64, int zc, const void *wall,
                    double offset, void *u, hipStream_t st);

}  // namespace gfmd
