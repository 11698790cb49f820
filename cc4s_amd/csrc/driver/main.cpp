// cc4s_ccsd: runs the CCSD steps of a cc4s operation file on one MI355X.
//
//   cc4s_ccsd [-i cc4s.in] [-o cc4s.out.yaml] [--dry-run]
//
// The operation file is the reference's (a YAML list of steps with name/in/out,
// Cc4s.cxx:52-175); the algorithms this drop-in executes are the CCSD path:
//   Read                    Read.cxx:27-43 (Reader.hpp: Tensor, TensorSet)
//   DefineHolesAndParticles DefineHolesAndParticles.cxx:38-96
//   SliceOperator           SliceOperator.cxx:50-131
//   VertexCoulombIntegrals  VertexCoulombIntegrals.cxx:36-52 (the field choice)
//   CoupledCluster          CoupledCluster.cxx:55-192, methods Ccsd and Drccd,
//                           DiisMixer,
//                           through the C ABI of include/cc4s_ccsd.h
//   Write                   Write.cxx:27-42 (Writer.hpp)
//   PerturbativeTriples(Reference)  PerturbativeTriplesReference.cxx:55-180
//   FiniteSizeCorrection    FiniteSizeCorrection.cxx:27-311 (complex field)
// Any other algorithm stops the run with "unknown algorithm: <name>", as the
// reference does for names it does not know.  Step arguments that are unquoted
// names of stored objects are replaced by the objects (Cc4s.cxx:177-195), the
// step's "out" names store its results (:197-211), and the report is
// rewritten after every step (:73-81) with in/out/statistics per step.
//
// Arguments of CoupledCluster beyond the reference's: precision (fp64 |
// mixed), mixedSwitch (1e-5), device (0).
#include "cc4s_ccsd.h"
#include "tensor_io.hpp"
#include "yaml.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace cc4s_amd;
using yaml::Node;
using yaml::NodePtr;

namespace {

struct Object {
  enum class Kind { Data, Tensor, TensorSet, SlicedVertex, Integrals };
  Kind kind = Kind::Data;
  NodePtr data;
  std::shared_ptr<HostTensor> tensor;
  std::shared_ptr<HostTensorSet> set;
  // SlicedVertex: the whole vertex; holes [0, No), particles [Np - Nv, Np)
  std::shared_ptr<HostTensor> vertex;
  int64_t No = 0, Nv = 0;
  // Integrals: field of the Coulomb integrals and the sliced vertex they use
  bool real_field = false;
  std::shared_ptr<Object> sliced_vertex;
  std::string label;
};
using ObjectPtr = std::shared_ptr<Object>;

std::string next_label() {
  static int counter = 0;
  char buf[32];
  std::snprintf(buf, sizeof(buf), "0x%06x", 0x1000 + 16 * ++counter);
  return buf;
}

ObjectPtr make_object(Object::Kind k) {
  auto o = std::make_shared<Object>();
  o->kind = k;
  o->label = next_label();
  return o;
}

struct Options {
  std::string in = "cc4s.in", out = "cc4s.out.yaml";
  bool dry_run = false;
};

class Runner {
 public:
  explicit Runner(const Options &o) : opt_(o) {}

  int run() {
    auto steps = yaml::parse_file(opt_.in);
    if (!steps->is_map()) throw std::runtime_error("expecting map as input");
    report_ = Node::map();
    auto env = Node::map();
    env->set("program", "cc4s_ccsd (MI355X CCSD drop-in)");
    env->set("dry run only", static_cast<int64_t>(opt_.dry_run ? 1 : 0));
    report_->set("executionEnvironment", env);
    executed_ = Node::map();
    report_->set("steps", executed_);
    const auto t0 = std::chrono::steady_clock::now();
    double total_flops = 0;
    for (size_t i = 0; i < steps->size(); ++i) {
      auto step = steps->get(i);
      if (!step || !step->is_map()) throw std::runtime_error("step " + std::to_string(i) + " is not a map");
      total_flops += run_step(i, step);
      emit();
    }
    const double total = seconds_since(t0);
    auto stats = Node::map();
    stats->set("realtime", total);
    stats->set("floatingPointOperations", total_flops);
    stats->set("flops", total > 0 ? total_flops / total : 0.0);
    report_->set("statistics", stats);
    emit();
    return 0;
  }

 private:
  static double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }

  void emit() { yaml::emit_file(report_, (opt_.dry_run ? "dry-" : "") + opt_.out); }

  // --- arguments --------------------------------------------------------------

  ObjectPtr symbol(const NodePtr &args, const std::string &key, bool required = true) {
    auto n = args ? args->get(key) : nullptr;
    if (!n) {
      if (required) throw std::runtime_error("expecting key '" + key + "'");
      return nullptr;
    }
    if (n->is_scalar() && !n->quoted) {
      auto it = storage_.find(n->value);
      if (it != storage_.end()) return it->second;
    }
    if (required) throw std::runtime_error("'" + key + "' does not name a stored object");
    return nullptr;
  }

  // the arguments as reported: stored objects by label
  NodePtr reported(const NodePtr &args) {
    auto r = Node::map();
    if (!args) return r;
    for (const auto &e : args->entries) {
      const auto &v = e.second;
      if (v && v->is_scalar() && !v->quoted && storage_.count(v->value))
        r->set(e.first, storage_.at(v->value)->label);
      else
        r->set(e.first, v);
    }
    return r;
  }

  double run_step(size_t i, const NodePtr &step) {
    const std::string name = yaml::get_text(step, "name", "");
    std::printf("step: %zu, %s\n", i + 1, name.c_str());
    std::fflush(stdout);
    NodePtr args = step->get("in");
    if (!args || !args->is_map()) args = Node::map();
    auto in_report = reported(args);
    std::map<std::string, ObjectPtr> results;
    NodePtr out_report = Node::map();
    double flops = 0;
    const auto t0 = std::chrono::steady_clock::now();
    if (name == "Read") {
      read(args, results);
    } else if (name == "DefineHolesAndParticles") {
      define_holes_and_particles(args, results);
    } else if (name == "SliceOperator") {
      slice_operator(args, results);
    } else if (name == "VertexCoulombIntegrals") {
      vertex_coulomb_integrals(args, results);
    } else if (name == "CoupledCluster") {
      flops = coupled_cluster(args, in_report, results, out_report);
    } else if (name == "Write") {
      write(args, out_report);
    } else if (name == "PerturbativeTriplesReference" || name == "PerturbativeTriples") {
      triples(args, results, name == "PerturbativeTriples");
    } else if (name == "FiniteSizeCorrection") {
      finite_size_correction(args, results);
    } else if (name == "SecondOrderPerturbationTheory") {
      flops = second_order(args, results);
    } else {
      throw std::runtime_error("unknown algorithm: " + name);
    }
    const double dt = seconds_since(t0);
    // store the results under the names of "out" (Cc4s.cxx:197-211)
    for (const auto &r : results)
      out_report->set(r.first, r.second->kind == Object::Kind::Data ? r.second->data
                                                                     : Node::scalar(r.second->label));
    auto outs = step->get("out");
    if (outs && outs->is_map()) {
      for (const auto &e : outs->entries) {
        auto it = results.find(e.first);
        if (it == results.end()) {
          std::fprintf(stderr, "warning: symbol '%s' should be stored but is not present in output\n",
                       e.first.c_str());
          continue;
        }
        storage_[e.second->text()] = it->second;
      }
    }
    auto rec = Node::map();
    rec->set("name", name);
    rec->set("in", in_report);
    rec->set("out", out_report);
    auto stats = Node::map();
    stats->set("realtime", dt);
    stats->set("floatingPointOperations", flops);
    stats->set("flops", dt > 0 ? flops / dt : 0.0);
    rec->set("statistics", stats);
    executed_->set(std::to_string(i), rec);
    std::printf("realtime %.6f s\n--\n", dt);
    std::fflush(stdout);
    return flops;
  }

  // --- algorithms ---------------------------------------------------------------

  void read(const NodePtr &args, std::map<std::string, ObjectPtr> &results) {
    const std::string file = yaml::get_text(args, "fileName", "");
    ReadItem item;
    try {
      item = read_file(file);
    } catch (const std::exception &e) {
      throw std::runtime_error("Failed to read file '" + file + "': " + e.what());
    }
    ObjectPtr o;
    if (item.kind == ReadItem::Kind::Tensor) {
      o = make_object(Object::Kind::Tensor);
      o->tensor = std::make_shared<HostTensor>(std::move(item.tensor));
    } else if (item.kind == ReadItem::Kind::TensorSet) {
      o = make_object(Object::Kind::TensorSet);
      o->set = std::make_shared<HostTensorSet>(std::move(item.set));
    } else {
      o = make_object(Object::Kind::Data);
      o->data = item.data;
    }
    results["destination"] = o;
  }

  void define_holes_and_particles(const NodePtr &args, std::map<std::string, ObjectPtr> &results) {
    auto eps = symbol(args, "eigenEnergies");
    if (eps->kind != Object::Kind::Tensor || eps->tensor->complex || eps->tensor->lens.size() != 1)
      throw std::runtime_error("expecting a Real64 tensor of rank 1 as 'eigenEnergies'");
    const auto &t = *eps->tensor;
    auto meta = t.meta_data;
    auto energies = meta ? meta->get("energies") : nullptr;
    if (!energies || !energies->is_map()) throw std::runtime_error("eigenEnergies without metaData.energies");
    const int64_t Np = static_cast<int64_t>(energies->size());
    const double fermi = yaml::get_real(meta, "fermiEnergy", 0.0);
    if (!meta->has("fermiEnergy")) throw std::runtime_error("eigenEnergies without metaData.fermiEnergy");
    int64_t No = 0;
    while (No < Np && energies->get(No)->real() < fermi) ++No;
    if (No == 0) throw std::runtime_error("Fermi energy below all eigen energies.");
    if (No >= Np) throw std::runtime_error("Fermi energy above all eigen energies.");
    if (t.lens[0] < Np) throw std::runtime_error("eigenEnergies shorter than metaData.energies");
    std::printf("number of holes     No: %lld\nnumber of particles Nv: %lld\nnumber of states    Np: %lld\n",
                static_cast<long long>(No), static_cast<long long>(Np - No),
                static_cast<long long>(Np));
    // eps_h = eps[0, No), eps_p = eps[No, Np) of the tensor elements (:65-83)
    HostTensor h, p;
    h.lens = {No};
    p.lens = {Np - No};
    h.dimension_types = p.dimension_types = t.dimension_types;
    h.unit = p.unit = t.unit;
    h.meta_data = p.meta_data = t.meta_data;
    h.data.assign(t.data.begin(), t.data.begin() + No);
    p.data.assign(t.data.begin() + No, t.data.begin() + Np);
    auto o = make_object(Object::Kind::TensorSet);
    o->set = std::make_shared<HostTensorSet>();
    (*o->set)["h"] = std::move(h);
    (*o->set)["p"] = std::move(p);
    results["slicedEigenEnergies"] = o;
  }

  static void sliced_counts(const ObjectPtr &e, int64_t &No, int64_t &Nv) {
    if (!e || e->kind != Object::Kind::TensorSet || !e->set->count("h") || !e->set->count("p"))
      throw std::runtime_error("expecting TensorSet of Real64 as 'slicedEigenEnergies'");
    No = e->set->at("h").lens.at(0);
    Nv = e->set->at("p").lens.at(0);
  }

  void slice_operator(const NodePtr &args, std::map<std::string, ObjectPtr> &results) {
    auto e = symbol(args, "slicedEigenEnergies");
    auto op = symbol(args, "operator");
    int64_t No, Nv;
    sliced_counts(e, No, Nv);
    if (op->kind != Object::Kind::Tensor) throw std::runtime_error("expecting a tensor as 'operator'");
    const auto &t = *op->tensor;
    if (t.lens.size() == 2 && t.dimension_types.size() == 2 && t.dimension_types[0] == "State" &&
        t.dimension_types[1] == "State") {
      results["slicedOperator"] = slice_state_operator(t, No, Nv);
      return;
    }
    // the CCSD path slices the vertex [G, State, State]
    if (t.lens.size() != 3 || t.dimension_types.size() != 3 || t.dimension_types[1] != "State" ||
        t.dimension_types[2] != "State" || t.dimension_types[0] == "State")
      throw std::runtime_error("expecting an operator [AuxiliaryField, State, State] to slice");
    if (t.lens[1] != t.lens[2] || t.lens[1] < No + Nv)
      throw std::runtime_error("operator has fewer states than holes + particles");
    if (!t.complex) throw std::runtime_error("expecting a Complex64 Coulomb vertex");
    std::printf("Slicing CoulombVertex into holes and particles.\n");
    auto o = make_object(Object::Kind::SlicedVertex);
    o->vertex = op->tensor;
    o->No = No;
    o->Nv = Nv;
    results["slicedOperator"] = o;
  }

  // a one-body operator [State, State] (the Fock operator) as the tensor set
  // {hh, hp, ph, pp}: holes [0, No), particles [Np - Nv, Np) along both
  // dimensions (SliceOperator.cxx:95-131)
  static ObjectPtr slice_state_operator(const HostTensor &t, int64_t No, int64_t Nv) {
    const int64_t Np = t.lens[0];
    if (t.lens[1] != Np || Np < No + Nv) throw std::runtime_error("operator has fewer states than holes + particles");
    std::printf("Slicing operator [State, State] into holes and particles.\n");
    const size_t es = t.complex ? 2 : 1;
    auto o = make_object(Object::Kind::TensorSet);
    o->set = std::make_shared<HostTensorSet>();
    const char *parts[4] = {"hh", "hp", "ph", "pp"};
    for (const char *part : parts) {
      int64_t begin[2], len[2];
      for (int d = 0; d < 2; ++d) {
        begin[d] = part[d] == 'h' ? 0 : Np - Nv;
        len[d] = part[d] == 'h' ? No : Nv;
      }
      HostTensor s;
      s.complex = t.complex;
      s.lens = {len[0], len[1]};
      s.dimension_types = t.dimension_types;
      s.unit = t.unit;
      s.meta_data = t.meta_data;
      s.data.resize(static_cast<size_t>(len[0] * len[1]) * es);
      for (int64_t y = 0; y < len[1]; ++y)
        for (int64_t x = 0; x < len[0]; ++x)
          for (size_t c = 0; c < es; ++c)
            s.data[es * (x + len[0] * y) + c] =
                t.data[es * ((begin[0] + x) + Np * (begin[1] + y)) + c];
      (*o->set)[part] = std::move(s);
    }
    return o;
  }

  void vertex_coulomb_integrals(const NodePtr &args, std::map<std::string, ObjectPtr> &results) {
    auto v = symbol(args, "slicedCoulombVertex");
    if (v->kind != Object::Kind::SlicedVertex) throw std::runtime_error("expecting a sliced Coulomb vertex");
    const bool half = yaml::get_boolean(v->vertex->meta_data, "halfGrid", false);
    std::printf("Using %s Coulomb integrals\n", half ? "real" : "complex");
    auto o = make_object(Object::Kind::Integrals);
    o->real_field = half;
    o->sliced_vertex = v;
    results["coulombIntegrals"] = o;
  }

  // Gamma [NG, No+Nv, No+Nv] -- the vertex itself, or its hole and particle
  // states when it has more states (SliceOperator.cxx:112-118) -- and
  // eps = [eps_h; eps_p]
  static const double *step_inputs(const ObjectPtr &vtx, const ObjectPtr &eps,
                                    std::vector<double> &compact, std::vector<double> &e) {
    const HostTensor &V = *vtx->vertex;
    const int64_t No = vtx->No, Nv = vtx->Nv;
    const int64_t NG = V.lens[0], NpV = V.lens[1], Np = No + Nv;
    const double *gamma = V.data.data();
    if (NpV != Np) {
      compact.resize(static_cast<size_t>(NG) * Np * Np * 2);
      auto state = [&](int64_t q) { return q < No ? q : NpV - Nv + (q - No); };
      for (int64_t q = 0; q < Np; ++q)
        for (int64_t r = 0; r < Np; ++r)
          std::memcpy(&compact[2 * NG * (r + Np * q)],
                      &V.data[2 * NG * (state(r) + NpV * state(q))], 2 * NG * sizeof(double));
      gamma = compact.data();
    }
    e.resize(Np);
    const auto &h = eps->set->at("h").data;
    const auto &pp = eps->set->at("p").data;
    std::copy(h.begin(), h.end(), e.begin());
    std::copy(pp.begin(), pp.end(), e.begin() + No);
    return gamma;
  }

  // PerturbativeTriplesReference / PerturbativeTriples: the (T) energy of the
  // amplitudes (PerturbativeTriplesReference.cxx:55-180; both names give the
  // same energy, PerturbativeTriples.cxx being the distributed variant)
  void triples(const NodePtr &args, std::map<std::string, ObjectPtr> &results, bool distributed) {
    auto ints = symbol(args, "coulombIntegrals");
    auto amps = symbol(args, "amplitudes");
    auto eps = symbol(args, "slicedEigenEnergies");
    if (ints->kind != Object::Kind::Integrals) throw std::runtime_error("expecting Coulomb integrals as 'coulombIntegrals'");
    if (yaml::get_integer(args, "cT", 0) != 0) throw std::runtime_error("cT (complete triples) is not supported");
    // (T*) of PerturbativeTriplesStar.hpp:34-103: the (T) energy times the sum
    // of the given MP2 pair energies over the finite-basis MP2 energy
    auto pairs = symbol(args, "mp2PairEnergies", false);
    if (args->has("mp2PairEnergies") && !pairs)
      throw std::runtime_error("'mp2PairEnergies' does not name a stored object");
    if (pairs && (pairs->kind != Object::Kind::Tensor || pairs->tensor->complex))
      throw std::runtime_error("expecting a Real64 tensor as 'mp2PairEnergies'");
    auto vtx = ints->sliced_vertex;
    int64_t No, Nv;
    sliced_counts(eps, No, Nv);
    if (amps->kind != Object::Kind::TensorSet || !amps->set->count("ph") || !amps->set->count("pphh"))
      throw std::runtime_error("unsupported tensor type as 'amplitudes'");
    const bool cplx = !ints->real_field;
    const HostTensor &t1 = amps->set->at("ph"), &t2 = amps->set->at("pphh");
    if (t1.complex != cplx || t2.complex != cplx || t1.elements() != Nv * No ||
        t2.elements() != Nv * Nv * No * No)
      throw std::runtime_error("amplitudes do not match the field or the sizes");
    std::vector<double> compact, e;
    const double *gamma = step_inputs(vtx, eps, compact, e);
    double et = 0.0, mp2 = 0.0;
    if (!opt_.dry_run) {
      cc4s_ccsd_params p;
      cc4s_ccsd_default_params(&p);
      p.field = cplx ? CC4S_FIELD_COMPLEX : CC4S_FIELD_REAL;
      p.No = No;
      p.Nv = Nv;
      p.NG = vtx->vertex->lens[0];
      p.max_residua = 1;
      p.device = static_cast<int>(yaml::get_integer(args, "device", 0));
      cc4s_ccsd_context *ctx = nullptr;
      auto check = [&](int rc) {
        if (rc != 0) {
          std::string msg = cc4s_ccsd_last_error(ctx);
          if (ctx) cc4s_ccsd_destroy(ctx);
          throw std::runtime_error(msg);
        }
      };
      check(cc4s_ccsd_create(&p, &ctx));
      check(cc4s_ccsd_set_inputs(ctx, gamma, e.data()));
      if (pairs) {
        // the first pass from no amplitudes: T2 = conj(V) / (e_i + e_j - e_a - e_b),
        // its energy 2 sum M_abij V_abij - M_abij V_abji is the finite-basis MP2 energy
        cc4s_ccsd_iteration it;
        check(cc4s_ccsd_iterate(ctx, &it));
        mp2 = it.energy;
      }
      check(cc4s_ccsd_set_amplitudes(ctx, t1.data.data(), t2.data.data()));
      check(cc4s_ccsd_triples(ctx, &et));
      cc4s_ccsd_destroy(ctx);
      std::printf("(T) correlation energy: %22.10f\n", et);
    }
    auto energy = Node::map();
    energy->set("correlation", et);
    auto star = [&]() {
      double cbs = 0.0;
      for (double x : pairs->tensor->data) cbs += x;
      const double v = mp2 != 0.0 ? cbs / mp2 * et : 0.0;
      std::printf("(T*)-Bsie energy correction: %23.15f\n", v);
      return v;
    };
    // key order as the two steps write them (PerturbativeTriples.cxx:238-254,
    // PerturbativeTriplesReference.cxx:201-213)
    if (pairs && distributed) energy->set("starCorrection", star());
    energy->set("unit", eps->set->at("h").unit);
    if (pairs && !distributed) energy->set("starCorrection", star());
    auto o = make_object(Object::Kind::Data);
    o->data = energy;
    results["energy"] = o;
  }

  // FiniteSizeCorrection: the transition structure factor on the momentum
  // grid at the amplitudes (FiniteSizeCorrection.cxx:49-133, on the GPU)
  // and its interpolation to the finite-size correction (:136-311, host)
  void finite_size_correction(const NodePtr &args, std::map<std::string, ObjectPtr> &results) {
    auto vtx = symbol(args, "slicedCoulombVertex");
    auto amps = symbol(args, "amplitudes");
    auto sv = symbol(args, "coulombVertexSingularVectors", false);
    auto pot = symbol(args, "coulombPotential");
    auto grid = symbol(args, "gridVectors");
    if (vtx->kind != Object::Kind::SlicedVertex) throw std::runtime_error("expecting a sliced Coulomb vertex");
    if (amps->kind != Object::Kind::TensorSet || !amps->set->count("ph") || !amps->set->count("pphh"))
      throw std::runtime_error("unsupported tensor type as 'amplitudes'");
    if (pot->kind != Object::Kind::Tensor || pot->tensor->complex || pot->tensor->lens.size() != 1)
      throw std::runtime_error("expecting a Real64 tensor [G] as 'coulombPotential'");
    if (grid->kind != Object::Kind::Tensor || grid->tensor->complex ||
        grid->tensor->lens.size() != 2 || grid->tensor->lens[0] != 3)
      throw std::runtime_error("expecting a Real64 tensor [3, G] as 'gridVectors'");
    const int64_t No = vtx->No, Nv = vtx->Nv, NF = vtx->vertex->lens[0];
    const int64_t NGg = grid->tensor->lens[1];
    if (pot->tensor->lens[0] != NGg) throw std::runtime_error("coulombPotential and gridVectors differ in G");
    const HostTensor &t1 = amps->set->at("ph"), &t2 = amps->set->at("pphh");
    if (t1.complex != t2.complex || t1.elements() != Nv * No || t2.elements() != Nv * Nv * No * No)
      throw std::runtime_error("the finite-size correction needs amplitudes of the vertex's sizes");
    // real amplitudes (real Coulomb integrals) are taken to the complex field,
    // the vertex being complex in either case (FiniteSizeCorrection.cxx:103-117)
    std::vector<double> c1, c2;
    const double *T1 = t1.data.data(), *T2 = t2.data.data();
    if (!t1.complex) {
      c1.assign(2 * t1.data.size(), 0.0);
      c2.assign(2 * t2.data.size(), 0.0);
      for (size_t k = 0; k < t1.data.size(); ++k) c1[2 * k] = t1.data[k];
      for (size_t k = 0; k < t2.data.size(); ++k) c2[2 * k] = t2.data[k];
      T1 = c1.data();
      T2 = c2.data();
    }
    const double *U = nullptr;
    if (sv) {
      const HostTensor &u = *sv->tensor;
      if (sv->kind != Object::Kind::Tensor || !u.complex || u.lens.size() != 2 || u.lens[0] != NGg ||
          u.lens[1] != NF)
        throw std::runtime_error("expecting Complex64 singular vectors [G, F] matching the grid and the vertex");
      U = u.data.data();
    } else if (NGg != NF) {
      throw std::runtime_error("without coulombVertexSingularVectors the grid must have the vertex's size");
    }
    // reciprocal lattice vectors from the grid's metaData Gi, Gj, Gk (:165-187)
    double B[9];
    const char *keys[3] = {"Gi", "Gj", "Gk"};
    for (int r = 0; r < 3; ++r) {
      auto v = grid->tensor->meta_data ? grid->tensor->meta_data->get(keys[r]) : nullptr;
      if (!v || v->size() != 3) throw std::runtime_error(std::string("gridVectors without metaData.") + keys[r]);
      for (int c = 0; c < 3; ++c) B[3 * r + c] = v->get(c)->real();
    }
    const int64_t N = yaml::get_integer(args, "interpolationGridSize", 20);
    std::vector<double> S(static_cast<size_t>(NGg), 0.0);
    double out[2] = {0.0, 0.0};
    if (!opt_.dry_run) {
      std::vector<double> compact, e;
      HostTensorSet eps;  // the structure factor does not use orbital energies
      eps["h"].data.assign(No, 0.0);
      eps["h"].lens = {No};
      eps["p"].data.assign(Nv, 0.0);
      eps["p"].lens = {Nv};
      auto eobj = make_object(Object::Kind::TensorSet);
      eobj->set = std::make_shared<HostTensorSet>(std::move(eps));
      const double *gamma = step_inputs(vtx, eobj, compact, e);
      cc4s_ccsd_params p;
      cc4s_ccsd_default_params(&p);
      p.field = CC4S_FIELD_COMPLEX;
      p.No = No;
      p.Nv = Nv;
      p.NG = NF;
      p.max_residua = 1;
      p.device = static_cast<int>(yaml::get_integer(args, "device", 0));
      cc4s_ccsd_context *ctx = nullptr;
      auto check = [&](int rc) {
        if (rc != 0) {
          std::string msg = cc4s_ccsd_last_error(ctx);
          if (ctx) cc4s_ccsd_destroy(ctx);
          throw std::runtime_error(msg);
        }
      };
      check(cc4s_ccsd_create(&p, &ctx));
      check(cc4s_ccsd_set_inputs(ctx, gamma, e.data()));
      check(cc4s_ccsd_set_amplitudes(ctx, T1, T2));
      check(cc4s_ccsd_transition_structure_factor(ctx, U, NGg, pot->tensor->data.data(), S.data()));
      cc4s_ccsd_destroy(ctx);
      if (cc4s_finite_size_correction(grid->tensor->data.data(), NGg, B, S.data(), N,
                                      vtx->vertex->unit, grid->tensor->unit, pot->tensor->unit,
                                      out) != 0)
        throw std::runtime_error(cc4s_ccsd_last_error(nullptr));
      std::printf("Finite-size energy correction:    %.10f\n", out[0] - out[1]);
    }
    auto sf = make_object(Object::Kind::Tensor);
    sf->tensor = std::make_shared<HostTensor>();
    sf->tensor->lens = {NGg};
    sf->tensor->dimension_types = {grid->tensor->dimension_types.size() > 1
                                       ? grid->tensor->dimension_types[1] : std::string("Momentum")};
    sf->tensor->data = std::move(S);
    results["transitionStructureFactor"] = sf;
    auto energy = Node::map();
    energy->set("correction", out[0] - out[1]);
    energy->set("uncorrected", out[1]);
    energy->set("corrected", out[0]);
    energy->set("unit", pot->tensor->unit * std::pow(grid->tensor->unit, 3.0));
    auto eo = make_object(Object::Kind::Data);
    eo->data = energy;
    results["energy"] = eo;
  }

  // spin degeneracy of SecondOrderPerturbationTheory.cxx:67-85: the Spin
  // property of the State dimension (propertyIndices.Spin of State.yaml in
  // the working directory, index -> value, TensorIo.cxx:30-63) marks an open
  // shell when fewer states carry the value 0 than there are states; closed
  // shell (2) otherwise and without the property
  static double spin_degeneracy() {
    if (!std::ifstream("State.yaml").good()) return 2.0;
    auto dim = yaml::parse_file("State.yaml");
    auto props = dim ? dim->get("propertyIndices") : nullptr;
    auto spin = props ? props->get("Spin") : nullptr;
    if (!spin || !spin->is_map()) return 2.0;
    size_t first = 0;
    for (const auto &e : spin->entries)
      if (e.second && e.second->integer() == 0) ++first;
    return first < spin->size() ? 1.0 : 2.0;
  }

  // SecondOrderPerturbationTheory (SecondOrderPerturbationTheory.cxx:54-149):
  // the second-order energy of the first-order amplitudes
  // T_ai = conj(f_ai) / (e_a - e_i), T_abij = conj(V_abij) / (e_a + e_b - e_i - e_j).
  // The doubles are the first pass of the iteration on the GPU (R2 = V_abij,
  // T2 = -R2 / D, the energy kernel): its direct term is -2 sum |V_abij|^2 / D
  // and its exchange term sum V_abij conj(V_baij) / D, the reference's
  // exchange sum V_abji conj(V_abij) / D by the pair symmetry V_abij = V_baji
  // of Coulomb integrals; the degeneracy g scales them by g^2 / 4 and g / 2.
  // The singles -g sum |f_ai|^2 / (e_a - e_i) are No * Nv host terms.
  double second_order(const NodePtr &args, std::map<std::string, ObjectPtr> &results) {
    auto ints = symbol(args, "coulombIntegrals");
    auto eps = symbol(args, "slicedEigenEnergies");
    auto fock = symbol(args, "slicedFockOperator", false);
    if (args->has("slicedFockOperator") && !fock)
      throw std::runtime_error("'slicedFockOperator' does not name a stored object");
    if (ints->kind != Object::Kind::Integrals)
      throw std::runtime_error("unsupported tensor type as 'coulombIntegrals'");
    int64_t No, Nv;
    sliced_counts(eps, No, Nv);
    auto vtx = ints->sliced_vertex;
    if (No != vtx->No || Nv != vtx->Nv) throw std::runtime_error("slices of different hole/particle counts");
    const bool cplx = !ints->real_field;
    const double g = spin_degeneracy();
    const auto &eh = eps->set->at("h").data;
    const auto &ep = eps->set->at("p").data;
    double singles = 0.0;
    if (fock) {
      if (fock->kind != Object::Kind::TensorSet || !fock->set->count("ph"))
        throw std::runtime_error("expecting a TensorSet {hh, hp, ph, pp} as 'slicedFockOperator'");
      const HostTensor &f = fock->set->at("ph");
      if (f.complex != cplx || f.lens.size() != 2 || f.lens[0] != Nv || f.lens[1] != No)
        throw std::runtime_error("slicedFockOperator.ph does not match the field or [particles, holes]");
      const size_t es = cplx ? 2 : 1;
      for (int64_t i = 0; i < No; ++i)
        for (int64_t a = 0; a < Nv; ++a) {
          const double *v = &f.data[es * (a + Nv * i)];
          const double m2 = v[0] * v[0] + (cplx ? v[1] * v[1] : 0.0);
          singles -= g * m2 / (ep[a] - eh[i]);
        }
    }
    std::printf("Contracting second order energy...\n");
    double direct = 0.0, exchange = 0.0, flops = 0.0;
    if (!opt_.dry_run) {
      std::vector<double> compact, e;
      const double *gamma = step_inputs(vtx, eps, compact, e);
      cc4s_ccsd_params p;
      cc4s_ccsd_default_params(&p);
      p.field = cplx ? CC4S_FIELD_COMPLEX : CC4S_FIELD_REAL;
      p.No = No;
      p.Nv = Nv;
      p.NG = vtx->vertex->lens[0];
      p.max_residua = 1;
      p.device = static_cast<int>(yaml::get_integer(args, "device", 0));
      cc4s_ccsd_context *ctx = nullptr;
      auto check = [&](int rc) {
        if (rc != 0) {
          std::string msg = cc4s_ccsd_last_error(ctx);
          if (ctx) cc4s_ccsd_destroy(ctx);
          throw std::runtime_error(msg);
        }
      };
      check(cc4s_ccsd_create(&p, &ctx));
      check(cc4s_ccsd_set_inputs(ctx, gamma, e.data()));
      cc4s_ccsd_iteration it;
      check(cc4s_ccsd_iterate(ctx, &it));
      cc4s_ccsd_destroy(ctx);
      direct = it.direct * g * g / 4.0;
      exchange = it.exchange * g / 2.0;
      flops = it.flops;
    }
    const double total = singles + direct + exchange;
    std::printf("correlation energy: %.12f\n  singles:  %.12f\n  direct:   %.12f\n  exchange: %.12f\n",
                total, singles, direct, exchange);
    auto energy = Node::map();
    energy->set("singles", singles);
    energy->set("direct", direct);
    energy->set("exchange", exchange);
    energy->set("correlation", total);
    energy->set("secondOrderSingles", singles);
    energy->set("secondOrderDirect", direct);
    energy->set("secondOrderExchange", exchange);
    energy->set("secondOrder", total);
    energy->set("unit", eps->set->at("h").unit);
    auto o = make_object(Object::Kind::Data);
    o->data = energy;
    results["energy"] = o;
    return flops;
  }

  double coupled_cluster(const NodePtr &args, NodePtr &in_report,
                         std::map<std::string, ObjectPtr> &results, NodePtr &out_report) {
    auto ints = symbol(args, "coulombIntegrals");
    auto eps = symbol(args, "slicedEigenEnergies");
    auto vtx = symbol(args, "slicedCoulombVertex");
    if (ints->kind != Object::Kind::Integrals) throw std::runtime_error("expecting Coulomb integrals as 'coulombIntegrals'");
    if (vtx->kind != Object::Kind::SlicedVertex) throw std::runtime_error("expecting a sliced Coulomb vertex");
    const std::string method = yaml::get_text(args, "method", "Ccsd");
    if (method != "Ccsd" && method != "Drccd")
      throw std::runtime_error("Unknown method: '" + method + "'");
    auto mixer = args->get("mixer");
    const std::string mixer_type = yaml::get_text(mixer, "type", "DiisMixer");
    if (mixer_type != "DiisMixer") throw std::runtime_error("Unknown mixer type: '" + mixer_type + "'");
    int64_t No, Nv;
    sliced_counts(eps, No, Nv);
    if (No != vtx->No || Nv != vtx->Nv) throw std::runtime_error("slices of different hole/particle counts");

    cc4s_ccsd_params p;
    cc4s_ccsd_default_params(&p);
    p.field = ints->real_field ? CC4S_FIELD_REAL : CC4S_FIELD_COMPLEX;
    p.No = No;
    p.Nv = Nv;
    p.NG = vtx->vertex->lens[0];
    p.max_iterations = yaml::get_integer(args, "maxIterations", p.max_iterations);
    p.energy_convergence = yaml::get_real(args, "energyConvergence", p.energy_convergence);
    p.amplitudes_convergence = yaml::get_real(args, "amplitudesConvergence", p.amplitudes_convergence);
    p.level_shift = yaml::get_real(args, "levelShift", p.level_shift);
    p.max_residua = yaml::get_integer(mixer, "maxResidua", p.max_residua);
    p.ppl = yaml::get_boolean(args, "ppl", true) ? 1 : 0;
    p.integrals_slice_size = yaml::get_integer(args, "integralsSliceSize", No);
    const std::string precision = yaml::get_text(args, "precision", "fp64");
    if (precision == "mixed")
      p.precision = CC4S_PRECISION_MIXED;
    else if (precision != "fp64")
      throw std::runtime_error("unknown precision '" + precision + "' (fp64 | mixed)");
    p.mixed_switch = yaml::get_real(args, "mixedSwitch", p.mixed_switch);
    const std::string ladder = yaml::get_text(args, "ladder", "paired");
    if (ladder == "fused")
      p.ladder = CC4S_LADDER_FUSED;
    else if (ladder == "paired")
      p.ladder = CC4S_LADDER_PAIRED;
    else if (ladder != "slab")
      throw std::runtime_error("unknown ladder '" + ladder + "' (slab | fused | paired)");
    p.device = static_cast<int>(yaml::get_integer(args, "device", 0));
    p.method = method == "Drccd" ? CC4S_METHOD_DRCCD : CC4S_METHOD_CCSD;
    p.linearized = yaml::get_boolean(args, "linearized", false) ? 1 : 0;
    p.adjacent_pairs_exchange = yaml::get_boolean(args, "adjacentPairsExchange", false) ? 1 : 0;
    // the report lists the defaults used, as the reference's argument map does
    in_report->set("method", method);
    in_report->set("maxIterations", p.max_iterations);
    in_report->set("energyConvergence", p.energy_convergence);
    in_report->set("amplitudesConvergence", p.amplitudes_convergence);
    in_report->set("levelShift", p.level_shift);
    if (p.method == CC4S_METHOD_DRCCD) {
      in_report->set("linearized", static_cast<int64_t>(p.linearized));
      in_report->set("adjacentPairsExchange", static_cast<int64_t>(p.adjacent_pairs_exchange));
    } else {
      in_report->set("ppl", static_cast<int64_t>(p.ppl));
      in_report->set("integralsSliceSize", p.integrals_slice_size);
    }

    ObjectPtr initial = symbol(args, "initialAmplitudes", false);
    if (args->has("initialAmplitudes") && !initial)
      throw std::runtime_error("'initialAmplitudes' does not name a stored object");

    const HostTensor &V = *vtx->vertex;
    const int64_t NG = V.lens[0];
    std::vector<double> compact, e;
    const double *gamma = step_inputs(vtx, eps, compact, e);

    if (opt_.dry_run) {
      std::printf("dry run: CCSD %s field, No=%lld Nv=%lld NG=%lld\n",
                  ints->real_field ? "real" : "complex", static_cast<long long>(No),
                  static_cast<long long>(Nv), static_cast<long long>(NG));
      out_report->set("dryRun", static_cast<int64_t>(1));
      return 0;
    }

    cc4s_ccsd_context *ctx = nullptr;
    auto check = [&](int rc) {
      if (rc != 0) {
        std::string msg = cc4s_ccsd_last_error(ctx);
        if (ctx) cc4s_ccsd_destroy(ctx);
        ctx = nullptr;
        throw std::runtime_error(msg);
      }
    };
    check(cc4s_ccsd_create(&p, &ctx));
    check(cc4s_ccsd_set_inputs(ctx, gamma, e.data()));
    const bool cplx = !ints->real_field;
    const size_t es = cplx ? 2 : 1;
    if (initial) {
      if (initial->kind != Object::Kind::TensorSet || !initial->set->count("ph") ||
          !initial->set->count("pphh"))
        throw std::runtime_error("expecting a TensorSet {ph, pphh} as 'initialAmplitudes'");
      const HostTensor &t1 = initial->set->at("ph"), &t2 = initial->set->at("pphh");
      if (t1.complex != cplx || t2.complex != cplx || t1.elements() != Nv * No ||
          t2.elements() != Nv * Nv * No * No)
        throw std::runtime_error("initialAmplitudes do not match the field or the sizes");
      check(cc4s_ccsd_set_amplitudes(ctx, t1.data.data(), t2.data.data()));
      std::printf("Using given initial amplitudes\n");
    }
    std::printf("Using method %s.\nUsing mixer DiisMixer. maxResidua: %lld\n", method.c_str(),
                static_cast<long long>(p.max_residua));
    std::printf("Maximum number of iterations: %lld\nUnless reaching energy convergence dE: %g\n"
                "and amplitudes convergence dR: %g\n",
                static_cast<long long>(p.max_iterations), p.energy_convergence,
                p.amplitudes_convergence);
    std::printf("Iter         Energy         dE           dR         time   TFLOP/s\n");
    std::vector<cc4s_ccsd_iteration> trace(static_cast<size_t>(std::max<int64_t>(1, p.max_iterations)));
    cc4s_ccsd_result r;
    check(cc4s_ccsd_run(ctx, &r, trace.data(), static_cast<int64_t>(trace.size())));
    double flops = 0;
    for (int64_t k = 0; k < r.iterations && k < static_cast<int64_t>(trace.size()); ++k) {
      const auto &it = trace[k];
      flops += it.flops;
      std::printf("%4lld %22.15f %12.5e %12.5e %8.3f %8.2f%s\n", static_cast<long long>(it.iteration),
                  it.energy, it.delta_energy, it.residuum_norm, it.seconds,
                  it.seconds > 0 ? it.flops / it.seconds / 1e12 : 0.0,
                  it.precision_used == CC4S_PRECISION_MIXED ? "  (mixed)" : "");
    }
    std::printf("\n%s correlation energy:          %.12f\n", method.c_str(), r.correlation);
    std::printf("2nd-order correlation energy:     %.12f\n", r.second_order);
    if (!r.convergence_reached) std::printf("A converged solution was not found\n");
    std::fflush(stdout);

    auto amps = make_object(Object::Kind::TensorSet);
    amps->set = std::make_shared<HostTensorSet>();
    HostTensor t1, t2;
    t1.complex = t2.complex = cplx;
    t1.lens = {Nv, No};
    t2.lens = {Nv, Nv, No, No};
    t1.data.resize(static_cast<size_t>(Nv * No) * es);
    t2.data.resize(static_cast<size_t>(Nv * Nv * No * No) * es);
    check(cc4s_ccsd_get_amplitudes(ctx, t1.data.data(), t2.data.data()));
    cc4s_ccsd_destroy(ctx);
    (*amps->set)["ph"] = std::move(t1);
    (*amps->set)["pphh"] = std::move(t2);

    // CoupledCluster.cxx:184-191,245-248; unit = unit(Gamma)^2
    auto energy = Node::map();
    energy->set("correlation", r.correlation);
    energy->set("direct", r.direct);
    energy->set("exchange", r.exchange);
    if (!std::isnan(r.second_order)) energy->set("secondOrder", r.second_order);
    energy->set("unit", V.unit * V.unit);
    results["amplitudes"] = amps;
    auto eobj = make_object(Object::Kind::Data);
    eobj->data = energy;
    results["energy"] = eobj;
    auto cobj = make_object(Object::Kind::Data);
    cobj->data = Node::scalar(r.convergence_reached ? "1" : "0");
    results["convergenceReached"] = cobj;
    return flops;
  }

  void write(const NodePtr &args, NodePtr &out_report) {
    auto src = symbol(args, "source");
    const std::string file = yaml::get_text(args, "fileName", "");
    const bool binary = yaml::get_boolean(args, "binary", false);
    NodePtr written;
    if (src->kind == Object::Kind::Tensor) {
      written = write_tensor_file(*src->tensor, file, binary);
    } else if (src->kind == Object::Kind::TensorSet) {
      written = write_tensor_set_file(*src->set, file, binary);
    } else if (src->kind == Object::Kind::Data) {
      yaml::emit_file(src->data, file);
      written = src->data;
    } else {
      throw std::runtime_error("Write supports tensors, tensor sets and data");
    }
    out_report->set("persitentSource", written);  // key spelled as Write.cxx:39
  }

  Options opt_;
  std::map<std::string, ObjectPtr> storage_;
  NodePtr report_, executed_;
};

}  // namespace

int main(int argc, char **argv) {
  Options opt;
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto value = [&]() -> std::string {
      if (a + 1 >= argc) throw std::runtime_error("missing value after " + s);
      return argv[++a];
    };
    try {
      if (s == "-i" || s == "--in")
        opt.in = value();
      else if (s == "-o" || s == "--out")
        opt.out = value();
      else if (s == "-l" || s == "--log")
        (void)value();  // logging goes to stdout
      else if (s == "-d" || s == "--dry-run")
        opt.dry_run = true;
      else if (s == "-h" || s == "--help") {
        std::printf("usage: cc4s_ccsd [-i cc4s.in] [-o cc4s.out.yaml] [--dry-run]\n");
        return 0;
      } else {
        throw std::runtime_error("unknown option " + s);
      }
    } catch (const std::exception &e) {
      std::fprintf(stderr, "cc4s_ccsd: %s\n", e.what());
      return 1;
    }
  }
  try {
    return Runner(opt).run();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    std::fflush(stdout);
    return 1;
  }
}
