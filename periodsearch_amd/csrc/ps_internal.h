// ps_internal.h -- declarations shared by the host (ps_host.cpp) and device (ps_device.hip) halves of
// libpsamd.so.  Not part of the public ABI.
#pragma once

// Records a message for ps_last_error() and returns code.
int ps_fail(int code, const char *fmt, ...);
