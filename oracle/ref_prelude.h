// Forced-include prelude for the compiled-reference checker build: pulls in
// the system headers first (include guards make the reference's own includes
// no-ops), then opens class access so the driver can read the set-up arrays.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>
#define private public
