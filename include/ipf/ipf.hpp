// ipf front end for MI355X: aggregate header. Provides the modules of the
// filtering path (the reference's include/ipf/ipf.hpp aggregates every module;
// this build carries core, parallel, filter, the pyramid functions of
// geometry, ssim of metrics, the filter-based callers of features and filter2DFFT).
#pragma once

#include "core.hpp"
#include "features.hpp"
#include "fft.hpp"
#include "filter.hpp"
#include "geometry.hpp"
#include "metrics.hpp"
#include "parallel.hpp"
