/*
 * visuals.h -- the quadruped's render geometry (visuals.cpp).
 */
#pragma once
#include "../../include/arc_world.h"

/* writes the robot's visuals (74: the torso's 10, then 16 per leg) into
 * out[0 .. cap) and returns how many it wrote */
int robot_visuals(arc_visual *out, int cap);
