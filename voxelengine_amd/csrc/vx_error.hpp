// Shared error reporting between the library's translation units: sets the
// message vx_last_error() returns on the calling thread and passes code through.
#pragma once
#include "vx/vx_api.h"

vx_status vx_set_error(vx_status code, const char* msg);
