"""The reference's headless benchmark gate metrics (headless_server.cpp:225-495), summarised.

The measurements themselves live in avbd/headless.py (the headlessBenchmark JSON line uses them):

* ground top: the highest top face of the static, axis-up box slabs at least 8 x 8 wide and at
  most 2 thick (isBenchmarkGroundReceiver, :384-393);
* ground clearance of each dynamic body: centre z - its vertical contact extent - ground top,
  zeroed within 1e-5 (:405-495); spheres' penetrations give the sphere-ground gate;
* final max speed: the largest |v| over all bodies (:283-286);
* final dynamic max height: the top of the dynamic bodies' bounding spheres, max(z + radius)
  (:311-330).
"""
from __future__ import annotations

from .headless import final_state_signature, ground_clearance, ground_contact_extent_z

contact_extent_z = ground_contact_extent_z


def pour_metrics(state) -> dict:
    sig = final_state_signature(state)
    gc = ground_clearance(state)
    out = {"dynamic_bodies": sig["dynamicBodies"], "final_max_speed": sig["maxSpeed"]}
    if gc["hasGround"]:
        out.update(ground_top=gc["groundTop"], max_sphere_ground_penetration=gc["maxPenetration"],
                   below_ground_spheres=gc["belowGroundSpheres"],
                   max_ground_body_penetration=gc["maxGroundBodyPenetration"])
    if sig["dynamicBodies"]:
        out["final_dynamic_max_height"] = sig["dynamicAabbMax"][2]
    return out
