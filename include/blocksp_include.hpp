// blocksp_include.hpp -- umbrella header of the MI355X drop-in (same name as the
// reference's umbrella, blocksp_include.hpp:1-16).  Link with -lblocksp_amd.
#pragma once
#include "blocksp/blocksp.hpp"
