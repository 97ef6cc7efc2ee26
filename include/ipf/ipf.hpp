// ipf front end for MI355X: aggregate header. Provides the modules of the
// filtering path (the reference's include/ipf/ipf.hpp aggregates every module;
// this build carries core, parallel, filter and the pyramid functions of
// geometry).
#pragma once

#include "core.hpp"
#include "filter.hpp"
#include "geometry.hpp"
#include "parallel.hpp"
