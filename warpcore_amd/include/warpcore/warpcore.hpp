// Umbrella header (include/warpcore/warpcore.cuh).
#pragma once

#include "base.hpp"
#include "status.hpp"
#include "tags.hpp"
#include "hashers.hpp"
#include "tile.hpp"
#include "fastmod.hpp"
#include "capacity.hpp"
#include "probing_schemes.hpp"
#include "storage.hpp"
#include "defaults.hpp"
#include "engine.hpp"
#include "single_value_hash_table.hpp"
#include "counting_hash_table.hpp"
#include "hash_set.hpp"
#include "random_distributions.hpp"
