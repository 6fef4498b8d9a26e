/* Forced include for compiling the reference's own sources with g++ on Linux
 * (build recipe in oracle/Makefile).  The reference builds with MSVC only:
 * its core.h defines M_PI<T> and INFINITY, which collide with the glibc
 * <math.h> macros, and src/main.cpp relies on a transitive <atomic>.  This
 * file pulls in the standard headers first and removes the two macros; it
 * supplies no declaration of the reference's own. */
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>
#include <math.h>
#undef M_PI
#undef INFINITY
