/* The oracle's walk, cell records and global coordinates compiled once more
 * with every double a dual number (orc_dual.hpp): forward-mode tangents of
 * particle positions and times with respect to the problem's parameters, in
 * up to ORC_ND directions given as the tangents of the problem's arrays
 * (dispersivities, diffusion, porosity).  Test infrastructure: the FP64
 * reference gradient against which the device's reverse sweep
 * (rwpt_adjoint) is checked; values equal the C oracle's bit for bit.
 * Entry points carry the prefix orcd_ and take orc_dual arrays
 * ((1 + ORC_ND) doubles per value) wherever the C oracle takes doubles. */
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "orc_dual.hpp"

#define ORC_DUAL_BUILD 1
#define orc_xoroshiro_next orcd_xoroshiro_next
#define orc_uniform orcd_uniform
#define orc_normal12 orcd_normal12
#define orc_rng_jump orcd_rng_jump
#define orc_rng_streams orcd_rng_streams
#define orc_advance orcd_advance
#define orc_cell_tables orcd_cell_tables
#define orc_global_xyz orcd_global_xyz
#define orc_histogram orcd_histogram
#define orc_newton_axis orcd_newton_axis
#define orc_quadratic_axis orcd_quadratic_axis
#define orc_kde_density orcd_kde_density
#define orc_kde_density_sel orcd_kde_density_sel
#define double orc_dual
extern "C" {
#include "rwpt_oracle.c"
}
#undef double
