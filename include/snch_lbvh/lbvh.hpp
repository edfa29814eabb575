// snch_lbvh/lbvh.hpp -- umbrella header, the counterpart of the reference's
// lbvh.cuh (lbvh.cuh:1-13): scenes, query descriptors, the OBJ loader, and
// the C ABI they sit on (libsnch.so, built for gfx950).
#pragma once

#include "../snch.h"
#include "scene.hpp"
#include "scene_loader.hpp"
