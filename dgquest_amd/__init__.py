"""dgquest_amd -- QuEST-API state-vector engine for one MI355X (gfx950).

* ``dgquest_amd.quest``    ctypes binding of the QuEST C API (libQuEST_f64/f32.so)
* ``dgquest_amd.engine``   runtime plumbing of the HIP engine (libdgq_engine.so)
* ``dgquest_amd.circuits`` seeded synthetic workloads
* ``dgquest_amd.build``    in-tree build of the libraries
"""
