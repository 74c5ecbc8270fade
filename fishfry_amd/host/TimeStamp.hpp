// (name, seconds) phase stamp -- the reference's TimeStamp (TimeStamp.hpp:7).
#pragma once

#include <string>
#include <utility>
#include <vector>

using TimeStamp = std::pair<std::string, double>;
