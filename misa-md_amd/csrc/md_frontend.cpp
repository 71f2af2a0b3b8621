// The reference frontend above the step: the YAML configuration
// (ConfigParser, frontend/config_parser.cpp:46-343, into ConfigValues,
// frontend/config_values.h:29-109 with the defaults of
// def_config_values.h and config_values.cpp:10-13,50-55) and the stage
// machine of MDSimulation (frontend/md_simulation.cpp:40-115) driving the
// resident step through the C-ABI.
//
// The stage machine runs the reference's per-step hooks at step
// boundaries, and hands the steps between two boundaries to one misa_step
// call, which runs them back to back on the device:
//   beforeStep(s): next stage when the current one is used up (its step
//     length, or def_timesteps_length when 0); the stage's recoil when
//     cur + 1 == collision_step and an origin dump path is set (dump, then
//     collisionStep); a rescale when cur % every_steps == 0;
//   postStep(s):  physical time += step length; a dump when
//     (s + 1) % atoms_dump_interval == 0; a thermo line when
//     (s + 1) % thermo_interval == 0.
#include <algorithm>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/misa_eam.h"
#include "yaml_lite.hpp"

namespace {

using misa::yaml::Node;

void copyStr(char* dst, size_t cap, const std::string& s) {
  const size_t n = std::min(cap - 1, s.size());
  std::memcpy(dst, s.data(), n);
  dst[n] = 0;
}

bool tooLong(const std::string& s) { return s.size() >= MISA_PATH_MAX; }

// ConfigParser::parseConfigSimulation .. parseStages, in the reference's
// order and with its messages.
bool parseSimulation(const Node* y, misa_config& c, std::string& e) {
  if (!y) return e = "\"simulation\" in config is not specified.", false;
  const Node* ps = y->get("phasespace");
  if (!ps || !ps->isSeq() || ps->size() != 3)
    return e = "\"simulation.phasespace\" must be a array with length 3.", false;
  for (int i = 0; i < 3; i++) c.phase_space[i] = static_cast<int64_t>(misa::yaml::asUInt(ps->at(i), 0));
  if (!y->get("cutoff_radius_factor")) return e = "cutoff_radius_factor must be set in config.", false;
  c.cutoff_radius_factor = misa::yaml::asDouble(y->get("cutoff_radius_factor"), 0.0);
  if (!y->get("lattice_const")) return e = "lattice const must be set in config.", false;
  c.lattice_const = misa::yaml::asDouble(y->get("lattice_const"), 0.0);
  c.time_step_length = misa::yaml::asDouble(y->get("def_timesteps_length"), 0.001);
  return true;
}

bool parseAlloy(const Node* y, misa_config& c, std::string& e) {
  if (!y) return e = "alloy config is not specified in creation mode.", false;
  c.alloy_create_seed = static_cast<int32_t>(misa::yaml::asInt(y->get("create_seed"), 1024));
  const Node* r = y->get("ratio");
  if (!r || !r->isMap()) return e = "alloy ratio must be a map.", false;
  c.alloy_ratio[0] = static_cast<int32_t>(misa::yaml::asInt(r->get("Fe"), 1));
  c.alloy_ratio[1] = static_cast<int32_t>(misa::yaml::asInt(r->get("Cu"), 0));
  c.alloy_ratio[2] = static_cast<int32_t>(misa::yaml::asInt(r->get("Ni"), 0));
  return true;
}

bool parseCreation(const Node* y, misa_config& c, std::string& e) {
  if (!y) return e = "\"creation\" in config is not specified.", false;
  c.create_phase = misa::yaml::asBool(y->get("create_phase"), true) ? 1 : 0;
  if (c.create_phase) {
    if (!y->get("create_t_set")) return e = "creation t_set must be specified.", false;
    c.create_t_set = misa::yaml::asDouble(y->get("create_t_set"), 0.0);
    c.create_seed = static_cast<int32_t>(misa::yaml::asInt(y->get("create_seed"), 1024));
  } else {
    const Node* f = y->get("read_phase_filename");
    if (!f) return e = "read phase file must be specified.", false;
    const std::string s = misa::yaml::asString(f, "");
    if (tooLong(s)) return e = "read_phase_filename is too long.", false;
    copyStr(c.read_phase_filename, MISA_PATH_MAX, s);
  }
  return parseAlloy(y->get("alloy"), c, e);
}

bool parsePotential(const Node* y, misa_config& c, std::string& e) {
  if (!y) return e = "\"potential\" in config is not specified.", false;
  if (!y->get("type")) return e = "potential file type must be specified.", false;
  const std::string t = misa::yaml::asString(y->get("type"), "");
  if (t.size() >= sizeof(c.potential_file_type)) return e = "potential file type is too long.", false;
  copyStr(c.potential_file_type, sizeof(c.potential_file_type), t);
  if (!y->get("file_path")) return e = "potential filename must be specified.", false;
  const std::string f = misa::yaml::asString(y->get("file_path"), "");
  if (tooLong(f)) return e = "potential filename is too long.", false;
  copyStr(c.potential_filename, MISA_PATH_MAX, f);
  return true;
}

bool parseOutput(const Node* y, misa_config& c, std::string& e) {
  if (!y) return e = "output section in config file is not specified.", false;
  const Node* d = y->get("dump");
  if (!d) return e = "dumping output section in config file is not specified.", false;
  if (d->get("atoms_dump_mode"))
    c.atoms_dump_mode = misa::yaml::asString(d->get("atoms_dump_mode"), "") == "copy" ? MISA_DUMP_COPY : MISA_DUMP_DEBUG;
  c.atoms_dump_interval = misa::yaml::asUInt(d->get("atoms_dump_interval"), UINT64_MAX);
  c.out_by_frame = misa::yaml::asBool(d->get("by_frame"), false) ? 1 : 0;
  const std::string path = misa::yaml::asString(d->get("atoms_dump_file_path"), "misa_mdl.out");
  const std::string origin = misa::yaml::asString(d->get("origin_dump_path"), "");
  if (tooLong(path) || tooLong(origin)) return e = "dump path is too long.", false;
  copyStr(c.atoms_dump_file_path, MISA_PATH_MAX, path);
  copyStr(c.origin_dump_path, MISA_PATH_MAX, origin);
  // the reference records this error and carries on parsing
  if (c.out_by_frame && path.find("{}") == std::string::npos) e = "error format of dump file path";
  const Node* t = y->get("thermo");
  if (!t) return e = "thermodynamics output section in config file is not specified.", false;
  c.thermo_interval = misa::yaml::asUInt(t->get("interval"), 0);
  const Node* l = y->get("logs");
  if (!l) return e = "output logs in config file is not specified.", false;
  if (misa::yaml::asString(l->get("logs_mode"), "console") == "console") {
    c.logs_mode = 0;
  } else {
    c.logs_mode = 1;
    std::string f = misa::yaml::asString(l->get("logs_filename"), "");
    if (f.empty()) f = "md.log";  // the reference derives a unique name from the cycle counter
    if (tooLong(f)) return e = "logs filename is too long.", false;
    copyStr(c.logs_filename, MISA_PATH_MAX, f);
  }
  return true;
}

bool parseStages(const Node* y, misa_config& c, std::string& e) {
  c.n_stages = 0;
  c.time_steps = 0;
  if (!y || !y->isSeq()) return true;  // as the reference: no stages is not a parse error
  if (y->size() > MISA_MAX_STAGES) return e = "too many stages.", false;
  for (size_t i = 0; i < y->size(); i++) {
    const Node* s = y->at(i);
    misa_stage& st = c.stages[c.n_stages];
    std::memset(&st, 0, sizeof(st));
    st.direction[0] = st.direction[1] = st.direction[2] = 1.0;
    st.rescale_every = 1;
    st.steps = misa::yaml::asUInt(s ? s->get("steps") : nullptr, 0);
    st.step_length = misa::yaml::asDouble(s ? s->get("step_length") : nullptr, 0.0);
    const Node* v = s ? s->get("set_v") : nullptr;
    if (v) {  // resolveConfigCollision
      st.collision_set = 1;
      st.collision_step = misa::yaml::asUInt(v->get("collision_step"), 0);
      const Node* lat = v->get("lat");
      if (!lat || !lat->isSeq() || lat->size() != 4)
        return e = "array length of value \"collision.lat\" must be 4.", false;
      for (int k = 0; k < 4; k++) st.collision_lat[k] = static_cast<int32_t>(misa::yaml::asInt(lat->at(k), 0));
      st.pka_energy = misa::yaml::asDouble(v->get("energy"), 0.0);
      const Node* dir = v->get("direction");
      if (!dir || !dir->isSeq() || dir->size() != 3)
        return e = "array length of value \"simulation.collision.collision_v\" must be 3.", false;
      for (int k = 0; k < 3; k++) st.direction[k] = misa::yaml::asDouble(dir->at(k), 0.0);
    }
    const Node* r = s ? s->get("rescale") : nullptr;
    if (r) {  // resolveConfigRescale
      st.rescale_set = 1;
      st.rescale_t = misa::yaml::asDouble(r->get("t"), 0.0);
      st.rescale_every = misa::yaml::asUInt(r->get("every_steps"), 0);
    }
    c.n_stages++;
    c.time_steps += st.steps;
  }
  return true;
}

void defaults(misa_config& c) {
  std::memset(&c, 0, sizeof(c));
  c.create_phase = 1;
  c.create_seed = 1024;
  c.alloy_create_seed = 1024;
  c.alloy_ratio[0] = 1;
  c.time_step_length = 0.001;
  c.atoms_dump_mode = MISA_DUMP_COPY;
  c.atoms_dump_interval = 1;
  copyStr(c.atoms_dump_file_path, MISA_PATH_MAX, "misa_mdl.out");
  copyStr(c.origin_dump_path, MISA_PATH_MAX, "origin_misa_mdl.out");
}

int parseText(const std::string& text, misa_config* out, char* err, int64_t err_cap) {
  std::string e;
  misa_config c;
  defaults(c);
  auto root = misa::yaml::parse(text, e);
  bool ok = root != nullptr;
  if (ok && !root->isMap()) {
    e = "the configuration is not a mapping";
    ok = false;
  }
  ok = ok && parseSimulation(root->get("simulation"), c, e) && parseCreation(root->get("creation"), c, e) &&
       parsePotential(root->get("potential"), c, e) && parseOutput(root->get("output"), c, e) &&
       parseStages(root->get("stages"), c, e);
  if (err && err_cap > 0) copyStr(err, static_cast<size_t>(err_cap), e);
  *out = c;
  return ok && e.empty() ? MISA_OK : MISA_E_ARG;
}

std::string formatPath(const std::string& pattern, uint64_t step) {
  // fmt::format(path, step) of a pattern with one "{}"
  const size_t k = pattern.find("{}");
  if (k == std::string::npos) return pattern;
  return pattern.substr(0, k) + std::to_string(step) + pattern.substr(k + 2);
}

struct Log {
  FILE* f = nullptr;
  bool own = false;
  void line(const char* fmt, ...) {
    if (!f) return;
    va_list ap;
    va_start(ap, fmt);
    std::vfprintf(f, fmt, ap);
    va_end(ap);
    std::fflush(f);
  }
  ~Log() {
    if (own && f) std::fclose(f);
  }
};

}  // namespace

extern "C" {

int misa_config_parse_text(const char* text, misa_config* out, char* err, int64_t err_cap) {
  if (!text || !out) return MISA_E_ARG;
  return parseText(text, out, err, err_cap);
}

int misa_config_parse(const char* path, misa_config* out, char* err, int64_t err_cap) {
  if (!path || !out) return MISA_E_ARG;
  std::ifstream in(path);
  if (!in.good()) {
    if (err && err_cap > 0) copyStr(err, static_cast<size_t>(err_cap), std::string("can not access the configure file: ") + path);
    return MISA_E_IO;
  }
  std::stringstream ss;
  ss << in.rdbuf();
  return parseText(ss.str(), out, err, err_cap);
}

int misa_md_simulate(const misa_config* c, int device, int32_t precision, misa_md_report* rep) {
  if (!c || c->n_stages < 0 || c->n_stages > MISA_MAX_STAGES) return MISA_E_ARG;
  if (!c->create_phase) return MISA_E_ARG;  // read mode (input::readPhaseSpace) is not part of this build
  if (std::string(c->potential_file_type) != "setfl") return MISA_E_ARG;
  misa_md_report r;
  std::memset(&r, 0, sizeof(r));
  Log log;
  if (c->logs_mode == 0) {
    log.f = stdout;
  } else {
    log.f = std::fopen(c->logs_filename, "a");
    log.own = log.f != nullptr;
    if (!log.f) return MISA_E_IO;
  }
  misa_geom g;
  for (int k = 0; k < 3; k++) g.phase_space[k] = c->phase_space[k];
  g.lattice_const = c->lattice_const;
  g.cutoff_radius_factor = c->cutoff_radius_factor;
  misa_handle* h = nullptr;
  int rc = misa_create(&g, device, &h);
  misa_dump* dump = nullptr;  // copy mode, one file for all frames
  auto finish = [&](int code) {
    if (dump) {
      int64_t n = 0;
      const int rc2 = misa_dump_close(dump, &n);
      r.dump_records += n;
      if (code == MISA_OK) code = rc2;
    }
    if (h) {
      misa_thermo t;
      if (code == MISA_OK && misa_thermo_compute(h, &t) == MISA_OK) r.thermo = t;
      misa_destroy(h);
    }
    if (rep) *rep = r;
    return code;
  };
  if (rc != MISA_OK) return finish(rc);
  const int32_t ratio[3] = {c->alloy_ratio[0], c->alloy_ratio[1], c->alloy_ratio[2]};
  if ((rc = misa_make_offsets(h)) != MISA_OK || (rc = misa_set_potential_file(h, c->potential_filename)) != MISA_OK ||
      (rc = misa_set_precision(h, precision)) != MISA_OK ||
      (rc = misa_build_world(h, static_cast<uint32_t>(c->create_seed), c->create_t_set, ratio,
                             static_cast<uint32_t>(c->alloy_create_seed))) != MISA_OK)
    return finish(rc);
  // prepareForStart: the first force evaluation with the first stage's step
  // length (the stage machine sets it again at step 0)
  double dt = c->n_stages > 0 && c->stages[0].step_length != 0.0 ? c->stages[0].step_length : c->time_step_length;
  if ((rc = misa_set_timestep(h, dt)) != MISA_OK || (rc = misa_compute_forces(h)) != MISA_OK) return finish(rc);

  // OutputDump::dump names its files by rank and step (output_dump.cpp:41-42)
  auto textDump = [&](uint64_t step) -> int {
    char name[64];
    std::snprintf(name, sizeof(name), "dump_0_%" PRIu64 ".atom", step);
    return misa_dump_text_state(h, name);
  };
  auto dumpFrame = [&](uint64_t step) -> int {
    if (c->atoms_dump_mode == MISA_DUMP_DEBUG) {
      r.dump_frames++;
      return textDump(step);
    }
    if (c->out_by_frame) {
      misa_dump* d = nullptr;
      int e = misa_dump_open(formatPath(c->atoms_dump_file_path, step).c_str(), &d);
      if (e != MISA_OK) return e;
      e = misa_dump_frame_state(d, h, step);
      int64_t n = 0;
      const int e2 = misa_dump_close(d, &n);
      r.dump_records += n;
      r.dump_frames++;
      return e != MISA_OK ? e : e2;
    }
    if (!dump) {
      const int e = misa_dump_open(c->atoms_dump_file_path, &dump);
      if (e != MISA_OK) return e;
    }
    r.dump_frames++;
    return misa_dump_frame_state(dump, h, step);
  };

  int next_stage = 0;
  misa_stage cur_stage;
  std::memset(&cur_stage, 0, sizeof(cur_stage));  // a stage of zero steps: step 0 moves to the first one
  uint64_t cur = 0;
  double phy_time = 0.0;
  const uint64_t total = c->time_steps;
  uint64_t step = 0;
  while (step < total) {
    // beforeStep(step)
    if (cur >= cur_stage.steps) {
      if (next_stage >= c->n_stages) {
        log.line("simulation: no more stages.\n");
        return finish(MISA_E_ARG);
      }
      cur_stage = c->stages[next_stage++];
      if (cur_stage.step_length == 0.0) cur_stage.step_length = c->time_step_length;
      cur = 0;
      dt = cur_stage.step_length;
      if ((rc = misa_set_timestep(h, dt)) != MISA_OK) return finish(rc);
      r.stages_run++;
    }
    if (cur_stage.collision_set && cur + 1 == cur_stage.collision_step && c->origin_dump_path[0]) {
      if (c->atoms_dump_mode == MISA_DUMP_DEBUG) {  // OutputDump::beforeCollision
        if ((rc = textDump(step)) != MISA_OK) return finish(rc);
      } else {  // OutputCopy::beforeCollision
        misa_dump* d = nullptr;
        if ((rc = misa_dump_open(c->origin_dump_path, &d)) != MISA_OK) return finish(rc);
        rc = misa_dump_frame_state(d, h, step);
        int64_t n = 0;
        const int rc2 = misa_dump_close(d, &n);
        if (rc != MISA_OK || rc2 != MISA_OK) return finish(rc != MISA_OK ? rc : rc2);
      }
      const int32_t lat[4] = {cur_stage.collision_lat[0], cur_stage.collision_lat[1], cur_stage.collision_lat[2],
                              cur_stage.collision_lat[3]};
      if ((rc = misa_collision_step(h, lat, cur_stage.direction, cur_stage.pka_energy)) != MISA_OK) return finish(rc);
      r.collisions++;
      log.line("simulation: collision at step %" PRIu64 ", %g eV\n", step, cur_stage.pka_energy);
    }
    if (cur_stage.rescale_set) {
      if (cur_stage.rescale_every == 0) return finish(MISA_E_ARG);  // cur % 0 in the reference
      if (cur % cur_stage.rescale_every == 0) {
        if ((rc = misa_rescale(h, cur_stage.rescale_t)) != MISA_OK) return finish(rc);
        r.rescales++;
      }
    }
    // steps until the next boundary at which a hook acts
    uint64_t n = std::min<uint64_t>(total - step, cur_stage.steps - cur);
    if (cur_stage.collision_set && c->origin_dump_path[0] && cur_stage.collision_step > cur + 1)
      n = std::min<uint64_t>(n, cur_stage.collision_step - 1 - cur);
    if (cur_stage.rescale_set) n = std::min<uint64_t>(n, cur_stage.rescale_every - cur % cur_stage.rescale_every);
    if (c->atoms_dump_interval > 0) n = std::min<uint64_t>(n, c->atoms_dump_interval - step % c->atoms_dump_interval);
    if (c->thermo_interval > 0) n = std::min<uint64_t>(n, c->thermo_interval - step % c->thermo_interval);
    n = std::max<uint64_t>(n, 1);
    if ((rc = misa_step(h, static_cast<int64_t>(n))) != MISA_OK) return finish(rc);
    step += n;
    cur += n;
    phy_time += static_cast<double>(n) * dt;
    // postStep(step - 1)
    if (c->atoms_dump_interval > 0 && step % c->atoms_dump_interval == 0) {
      if ((rc = dumpFrame(step)) != MISA_OK) return finish(rc);
    }
    if (c->thermo_interval > 0 && step % c->thermo_interval == 0) {
      misa_thermo t;
      if ((rc = misa_thermo_compute(h, &t)) != MISA_OK) return finish(rc);
      log.line("energy: kinetic energy = %g, T = %g.\n", t.kinetic, t.temperature);
      r.thermo_lines++;
    }
  }
  r.steps = step;
  r.phy_time = phy_time;
  log.line("simulation: %" PRIu64 " steps, %g ps.\n", step, phy_time);
  return finish(MISA_OK);
}

}  // extern "C"
