// The library's bookkeeping entry points: version, errors, the constant layout,
// the compiled mechanisms and the case handle.
#include <cstdlib>
#include <cstring>

#include "generated/mechanisms.hpp"
#include "pga_common.hpp"
#include "pga_internal.hpp"

namespace pga {

static thread_local char errbuf[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(errbuf, sizeof errbuf, fmt, ap);
  va_end(ap);
}

struct MechInfo {
  const char *name, *signature;
  int ns, table;
};

// the compiled mechanisms, after a sentinel so a library built for an inert case (none) still has
// an array
static const MechInfo mechs_[] = {
    {"", "", 0, 0},
#define X(NAME, IDX) {Mech_##NAME::NAME_, Mech_##NAME::SIGNATURE, Mech_##NAME::NS, Mech_##NAME::TABLE},
    PGA_MECHANISMS(X)
#undef X
};
static const MechInfo *const mechs = mechs_ + 1;
static const int nmechs = sizeof mechs_ / sizeof mechs_[0] - 1;

}  // namespace pga

using namespace pga;

extern "C" int32_t pga_abi_version(void) { return PGA_ABI_VERSION; }

extern "C" const char *pga_last_error(void) { return errbuf; }

extern "C" int32_t pga_const_offset(int32_t ns, const char *name, int32_t *count) {
  if (ns < 1 || ns > MAXNS || !name) return -1;
  const Offsets o = offsets(ns);
  struct {
    const char *n;
    int off, cnt;
  } rows[] = {
      {"MW", o.MW, ns},           {"MWinv", o.MWinv, ns},     {"Rsp", o.Rsp, ns},   {"cppolyR", o.cppolyR, ns * NT},
      {"h298R", o.h298R, ns},     {"cp0", o.cp0, ns},         {"gpoly", o.gpoly, ns * NG}, {"h298", o.h298, ns},
      {"mu0", o.mu0, ns},         {"kappa0", o.kappa0, ns},   {"lewis", o.lewis, ns},
      {"hpoly", o.hpoly, ns * NT}, {"MWqinv", o.MWqinv, ns},  {"wilkepair", o.wilkepair, ns * ns},
      {"mupoly", o.mupoly, ns * NT}, {"kappapoly", o.kappapoly, ns * NT}, {"cppoly", o.cppoly, ns * NT},
      {"cubic", o.cubic, 4},       {"cubicsqrtac", o.cubicsqrtac, ns}, {"cubicb", o.cubicb, ns},
      {"cubicda", o.cubicda, ns},  {"cubicmTc", o.cubicmTc, ns}, {"cubiconepm", o.cubiconepm, ns},
      {"cubicsqrtacMW", o.cubicsqrtacMW, ns}, {"cubicbMW", o.cubicbMW, ns}, {"vshift", o.vshift, ns},
      {"vshiftMW", o.vshiftMW, ns}, {"cubicdaMW", o.cubicdaMW, ns}, {"acentric", o.acentric, ns},
      {"chungsigma", o.chungsigma, ns}, {"chungepsk", o.chungepsk, ns}, {"chungvisc", o.chungvisc, 40},
      {"chungcond", o.chungcond, 28}, {"pairfit", o.pairfit, ns * ns * NT}, {"selffit", o.selffit, ns * NT},
      {"bathof", o.bathof, ns * ns}, {"diffmeta", o.diffmeta, 2},
  };
  for (auto &r : rows)
    if (!strcmp(r.n, name)) {
      if (count) *count = r.cnt;
      return r.off;
    }
  return -1;
}

extern "C" int32_t pga_const_size(int32_t ns) { return (ns < 1 || ns > MAXNS) ? -1 : offsets(ns).size; }

extern "C" int32_t pga_mechanism_count(void) { return nmechs; }

extern "C" const char *pga_mechanism_signature(int32_t i) { return (i >= 0 && i < nmechs) ? mechs[i].signature : ""; }

extern "C" const char *pga_mechanism_name(int32_t i) { return (i >= 0 && i < nmechs) ? mechs[i].name : ""; }

extern "C" int32_t pga_mechanism_table_size(int32_t i) { return (i >= 0 && i < nmechs) ? mechs[i].table : -1; }

extern "C" int32_t pga_mechanism_find(const char *sig) {
  if (!sig) return -1;
  for (int i = 0; i < nmechs; i++)
    if (!strcmp(mechs[i].signature, sig)) return i;
  return -1;
}

extern "C" int32_t pga_case_create(const pga_case_desc *desc, const double *dev_consts, const double *dev_chem,
                                   pga_case **out) {
  if (!desc || !dev_consts || !out) {
    set_error("pga_case_create: null argument");
    return PGA_EINVAL;
  }
  if (desc->ns < 1 || desc->ns > MAXNS) {
    set_error("pga_case_create: %d species; 1 to %d", desc->ns, MAXNS);
    return PGA_EINVAL;
  }
  if (desc->eos < 0 || desc->eos > 2) {
    set_error("pga_case_create: no eos %d (0 cpg, 1 tpg, 2 cubic real gas)", desc->eos);
    return PGA_EUNSUP;
  }
  if (desc->eos == 2 && !realgas_compiled(desc->ns)) {
    set_error("pga_case_create: the real gas is not compiled for %d species here (a library built for the case "
              "has it)", desc->ns);
    return PGA_EUNSUP;
  }
  if (desc->trans < -1 || desc->trans > 2 || (desc->trans == 2 && desc->eos != 2)) {
    set_error("pga_case_create: no transport model %d (-1 none, 0 constant, 1 kinetic theory, 2 Chung on a real gas)",
              desc->trans);
    return PGA_EUNSUP;
  }
  if (desc->diffusion < 0 || desc->diffusion > 1) {
    set_error("pga_case_create: no diffusion model %d (0 Lewis numbers, 1 binary)", desc->diffusion);
    return PGA_EUNSUP;
  }
  if (desc->mechanism >= 0) {
    if (desc->mechanism >= nmechs || !dev_chem) {
      set_error("pga_case_create: mechanism %d not compiled, or no table", desc->mechanism);
      return PGA_EUNSUP;
    }
    if (mechs[desc->mechanism].ns != desc->ns) {
      set_error("pga_case_create: mechanism %s has %d species, the case %d", mechs[desc->mechanism].name,
                mechs[desc->mechanism].ns, desc->ns);
      return PGA_EINVAL;
    }
  }
  pga_case *c = static_cast<pga_case *>(std::calloc(1, sizeof(pga_case)));
  if (!c) {
    set_error("pga_case_create: out of host memory");
    return PGA_ERANGE;
  }
  c->desc = *desc;
  c->K = dev_consts;
  c->chem = desc->mechanism >= 0 ? dev_chem : nullptr;
  *out = c;
  return PGA_OK;
}

extern "C" int32_t pga_case_destroy(pga_case *c) {
  std::free(c);
  return PGA_OK;
}
