// vx3mi_vxa.cpp -- VXA/VXD reader (placeholder until the XML reader lands).
#include "vx3mi_internal.h"

using namespace vx3mi;

extern "C" {
int vx3mi_model_from_xml(const char *, const char *, vx3mi_model **out) {
    if (out) *out = nullptr;
    set_error("VXA reader not built");
    return VX3MI_E_PARSE;
}
int vx3mi_model_from_vxa(const char *, const char *, vx3mi_model **out) {
    if (out) *out = nullptr;
    set_error("VXA reader not built");
    return VX3MI_E_PARSE;
}
}
