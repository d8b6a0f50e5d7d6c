// Reference header name (hamvec.h) for plugin code written against the
// reference: the same classes and signatures, from hamx_host.h.
#pragma once

#include "hamx_compat.h"
