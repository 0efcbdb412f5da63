// Host-only stand-in for the runtime header of generated programs, for the
// front end's CPU tests of programs that make no runtime calls.
#pragma once
#include <cstring>
#include <stdexcept>
#include <string>
