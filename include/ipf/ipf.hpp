// ipf front end for MI355X: aggregate header. Provides the modules of the
// filtering path (the reference's include/ipf/ipf.hpp aggregates every module;
// this build carries core, parallel, filter, the pyramid functions of
// geometry, ssim of metrics and the filter-based callers of features).
#pragma once

#include "core.hpp"
#include "features.hpp"
#include "filter.hpp"
#include "geometry.hpp"
#include "metrics.hpp"
#include "parallel.hpp"
