"""MI355X tree gravity: device-resident Morton keys, octree, node moments and softened tree walks
behind a C ABI (include/gz_grav.h), with a host mirror of the shmem tree call surface."""
