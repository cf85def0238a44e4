/*
 * setup.h -- everything derived once per simulation on the host: the robot
 * built in binary32, the Model constants of the step kernel, the reset and
 * template constants, and the contact channel of each sphere
 * (csrc/arc_sim.c:26-54).  The device library and the host-compiled parity
 * harness build it with the same function.
 */
#pragma once
#include <vector>
#include "device_state.h"
#include "model.h"

namespace xpbd {

struct SimSetup {
    int grid;
    HostWorld tmpl;
    RobotInfo info;
    Model model;
    ResetK rk;
    std::vector<uint32_t> jtab; /* rk's jump tables on the host (rk.jtab points here) */
    TemplateK tk;
    int channel[MAX_SPHERES];
};

/* false (with a message on stderr) when the arguments are invalid or the
 * built robot does not have the topology the engine is compiled for */
bool sim_setup(SimSetup &S, double dt, int substeps, int grid, double cell);

} /* namespace xpbd */
