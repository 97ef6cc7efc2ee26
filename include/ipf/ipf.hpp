// ipf front end for MI355X: aggregate header. Provides the modules of the
// filtering path (the reference's include/ipf/ipf.hpp aggregates every module;
// this build carries core, parallel, filter, the pyramid functions of
// geometry, ssim of metrics, the filter-based callers of features, filter2DFFT
// and the Netpbm / PFM codecs of io; npy.hpp adds .npy arrays).
#pragma once

#include "core.hpp"
#include "features.hpp"
#include "fft.hpp"
#include "filter.hpp"
#include "geometry.hpp"
#include "io.hpp"
#include "metrics.hpp"
#include "parallel.hpp"
