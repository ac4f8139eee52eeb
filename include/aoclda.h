/*
 * Umbrella header of the MI355X build of AOCL-DA's distance-bound path
 * (source/include/aoclda.h of the reference includes every chapter; this one includes the
 * chapters this library implements plus the GPU extension header).
 */
#ifndef AOCLDA_H
#define AOCLDA_H

#include "aoclda_approximate_neighbors.h"
#include "aoclda_dbscan.h"
#include "aoclda_error.h"
#include "aoclda_handle.h"
#include "aoclda_kmeans.h"
#include "aoclda_metrics.h"
#include "aoclda_nearest_neighbors.h"
#include "aoclda_options.h"
#include "aoclda_result.h"
#include "aoclda_types.h"
#include "aoclda_utils.h"
#include "aoclda_amd.h"

#endif
