#pragma once
#include "vec3.hpp"
#include "quat.hpp"
#include "mat3.hpp"
