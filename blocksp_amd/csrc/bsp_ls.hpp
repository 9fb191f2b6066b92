// bsp_ls.hpp -- the device view (rs::LSState) of a GMRES least-squares state (DeviceLS).
#pragma once

#include "bsp_internal.hpp"
#include "bsp_restart.hpp"

namespace bsp {

inline rs::LSState ls_state_of(const DeviceLS& ls) {
    return {ls.H.as<double>(),   ls.R.as<double>(),   ls.cs.as<double>(), ls.sn.as<double>(),
            ls.g.as<double>(),   ls.Q0t.as<double>(), ls.y.as<double>(),  ls.res.as<double>(),
            ls.meta.as<int>(),   ls.ld,               ls.maxm};
}

}  // namespace bsp
