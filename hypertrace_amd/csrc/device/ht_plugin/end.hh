// Device plugin region, part 2 of 2: closes what ht_plugin/begin.hh opened.
// The global vocabulary (types, bridged functions, ht_plugin:: adaptors)
// stays; the macros that would rewrite later code are withdrawn.
#pragma once

#ifndef HT_PLUGIN_REGION
#error "ht_plugin/end.hh without ht_plugin/begin.hh"
#endif

#pragma clang force_cuda_host_device end

#undef __global
#undef __allow_unused_params__
#undef __allow_multiple__
#undef float3
#undef double3
#undef make
#undef $2
#undef $3
#undef define_shape_interface
#undef define_object_interface
#undef define_material_interface
#undef define_background_interface
#undef define_view_interface
