// The classes of hamx_host.h under the reference's global names, so that
// plugin code written against the reference's headers (bfield.h,
// tefield.h, crefield.h, hamvec.h, hamtype.h, param.h, grid.h,
// integrator.h, pipeline.h) compiles unchanged against this build: put
// this directory first on the include path.
#pragma once

#include "hamx_host.h"

using hamx_host::ham_float;
using hamx_host::ham_int;
using hamx_host::ham_sint;
using hamx_host::ham_uint;
using hamx_host::Hamvec;

using hamx_host::Param;

using hamx_host::Grid;
using hamx_host::Grid_breg;
using hamx_host::Grid_brnd;
using hamx_host::Grid_cre;
using hamx_host::Grid_obs;
using hamx_host::Grid_tereg;
using hamx_host::Grid_ternd;
using hamx_host::Hampix;

using hamx_host::Breg;
using hamx_host::Breg_cart;
using hamx_host::Breg_helix;
using hamx_host::Breg_jaffe;
using hamx_host::Breg_jf12;
using hamx_host::Breg_lsa;
using hamx_host::Breg_unif;
using hamx_host::Brnd;
using hamx_host::Brnd_es;
using hamx_host::Brnd_jf12;
using hamx_host::Brnd_mhd;
using hamx_host::TEreg;
using hamx_host::TEreg_unif;
using hamx_host::TEreg_ymw16;
using hamx_host::TErnd;
using hamx_host::TErnd_dft;
using hamx_host::CREfield;
using hamx_host::CRE_ana;
using hamx_host::CRE_num;
using hamx_host::CRE_unif;

using hamx_host::Integrator;
using hamx_host::Pipeline;
