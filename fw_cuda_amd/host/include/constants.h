// constants.h -- drop-in for the reference's include/constants.h:3-9.
#pragma once

constexpr int INF = 1'000'000'000;
constexpr int MAX_N = 32769;
constexpr int MAX_EDGE_WEIGHT = 25'000;
using WeightType = int;
