// cgfd_internal.h -- shared by the library's host files, not part of the ABI.
#pragma once

// set the calling thread's cgfd_last_error() message
void cgfd_internal_set_error(const char *msg);
