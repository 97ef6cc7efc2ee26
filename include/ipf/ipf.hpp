// ipf front end for MI355X: aggregate header. Provides the modules of the
// filtering path (the reference's include/ipf/ipf.hpp aggregates every module;
// this build carries core, parallel and filter).
#pragma once

#include "core.hpp"
#include "filter.hpp"
#include "parallel.hpp"
