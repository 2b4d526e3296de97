"""MI355X hybrid pass of VulkanHybridRaytracer: GPU LBVH, shadow and
reflection traversal and deferred PBR shading behind the C ABI of
include/hr/hr.h (libhr.so, HIP for gfx950)."""
