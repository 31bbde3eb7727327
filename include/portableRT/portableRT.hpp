// Umbrella header of the portableRT API with the MI355X backend (portableRT.hpp in the
// reference, v0.3.3).
#pragma once
#include "backend.hpp"
#include "core.hpp"
#include "hitreg.hpp"
#include "intersect_mi355x.hpp"
#include "nearest_hits_impl.hpp"
