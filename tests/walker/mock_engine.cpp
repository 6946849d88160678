// mock_engine.cpp -- a CPU engine for the program walker (test infrastructure).
//
// Implements the Engine interface with host vectors, including device-side
// predicates (predicate = popcount of the condition, select by its value), so
// the walker's if-conversion can be checked on the CPU: the same program run
// with predicates on and off must give the same outputs, op count and
// variable count.  Built by tests/test_walker_ifconv_cpu.py with g++ (and the
// address and undefined-behaviour sanitizers); never part of the library.
//
//   mock_engine <prog.bs> <case.bsdata> <1 predicates | 0 host decisions>
#include "executor.h"
#include <cstdio>
#include <fstream>
#include <sstream>
using namespace bsz;
struct MV : StreamVal { std::vector<uint64_t> w; };
struct MP : PopCell {};
class Mock final : public Engine {
public:
  Geom g; uint64_t nw=1;
  std::shared_ptr<MV> mk() { auto v=std::make_shared<MV>(); v->w.assign(nw,0); return v; }
  static MV* M(const SV& s){ auto* p=dynamic_cast<MV*>(s.get()); if(!p){fprintf(stderr,"not MV\n"); abort();} return p; }
  void mask(std::vector<uint64_t>& w){ if(g.bitlength>0){ for(uint64_t i=g.lw;i<nw;++i)w[i]=0; w[g.lw-1]&=g.lastmask; } }
  void begin(const Geom& gg) override { g=gg; nw=gg.lw; }
  SV constant(bool ones) override { auto v=mk(); v->konst = ones; if(ones){ for(auto& x:v->w)x=~0ull; mask(v->w);} return v; }
  SV input(const uint64_t* h, uint64_t n) override { auto v=mk(); for(uint64_t i=0;i<nw&&i<n;++i)v->w[i]=h[i]; mask(v->w); return v; }
  SV op_not(const SV& a) override { auto v=mk(); for(uint64_t i=0;i<nw;++i)v->w[i]=~M(a)->w[i]; mask(v->w); return v; }
  SV op_bin(BinOp op, const SV& a, const SV& b) override { auto v=mk(); auto&x=M(a)->w; auto&y=M(b)->w; unsigned __int128 c=0;
    for(uint64_t i=0;i<nw;++i){ switch(op){case BinOp::And:v->w[i]=x[i]&y[i];break;case BinOp::Or:v->w[i]=x[i]|y[i];break;case BinOp::Xor:v->w[i]=x[i]^y[i];break;
      case BinOp::Add:{unsigned __int128 s=(unsigned __int128)x[i]+y[i]+c; v->w[i]=(uint64_t)s; c=s>>64; break;}}}
    mask(v->w); return v; }
  SV op_shift(bool left, const SV& a, uint64_t k) override { auto v=mk(); auto&x=M(a)->w; uint64_t q=k/64; unsigned b=k%64;
    for(uint64_t i=0;i<nw;++i){ if(left){ uint64_t hi= i>=q? x[i-q]:0, lo= (b&&i>=q+1)? x[i-q-1]:0; v->w[i]= b? (hi<<b)|(lo>>(64-b)) : hi; }
      else { uint64_t lo= i+q<nw? x[i+q]:0, hi=(b&&i+q+1<nw)? x[i+q+1]:0; v->w[i]= b? (lo>>b)|(hi<<(64-b)):lo; } }
    mask(v->w); return v; }
  PopRef popcount(const SV& a) override { auto p=std::make_shared<MP>(); int64_t t=0; if(g.bitlength>0) for(uint64_t i=0;i<g.lw;++i) t+=__builtin_popcountll(M(a)->w[i]); p->value=t; p->ready=true; return p; }
  int64_t resolve(const PopRef& p) override { return p->value; }
  bool is_nonzero(const SV& a) override { if(g.bitlength<=0) return false; for(uint64_t i=0;i<g.lw;++i) if(M(a)->w[i]) return true; return false; }
  void download(const SV& a, uint64_t* h) override { for(uint64_t i=0;i<g.lw;++i) h[i]=M(a)->w[i]; }
  void flush() override {}
  void sync() override {}
  int64_t n_kernel_launches() const override { return 0; }
  int64_t n_mallocs() const override { return 0; }
  size_t peak_bytes() const override { return 0; }
  hipStream_t stream() const override { return nullptr; }
  bool preds = true;
  bool supports_predicates() const override { return preds; }
  PopRef predicate(const SV& c) override { return popcount(c); }
  SV select(const PopRef& p, const SV& t, const SV& f) override { if (t==f) return t; auto v=mk(); v->w = p->value? M(t)->w : M(f)->w; return v; }
};
int main(int argc, char** argv) {
  std::ifstream f(argv[1]); std::stringstream ss; ss<<f.rdbuf();
  auto prog = parse_program(ss.str());
  auto cases = read_bsdata_file(argv[2]);
  Mock m;
  if (argc > 3) m.preds = argv[3][0] == '1';
  RunResult r = execute_program(*prog, cases[0], m);
  auto hex = [](const std::vector<uint64_t>& w) {
    std::string o = "0x";
    char b[20];
    for (size_t i = w.size(); i-- > 0;) { std::snprintf(b, sizeof b, "%016llx", (unsigned long long)w[i]); o += b; }
    return o;
  };
  printf("{\"op_count\":%lld,\"n_variables\":%lld,\"streams\":{", (long long)r.op_count, (long long)r.n_variables);
  bool first = true;
  for (auto& kv : r.streams) { printf("%s\"%s\":\"%s\"", first ? "" : ",", kv.first.c_str(), hex(kv.second).c_str()); first = false; }
  printf("},\"arrays\":{");
  first = true;
  for (auto& kv : r.arrays) {
    printf("%s\"%s\":{", first ? "" : ",", kv.first.c_str()); first = false;
    bool f2 = true;
    for (auto& e : kv.second) { printf("%s\"%lld\":\"%s\"", f2 ? "" : ",", (long long)e.first, hex(e.second).c_str()); f2 = false; }
    printf("}");
  }
  printf("},\"ints\":{");
  first = true;
  for (auto& kv : r.ints) { printf("%s\"%s\":%lld", first ? "" : ",", kv.first.c_str(), (long long)kv.second); first = false; }
  printf("}}\n");
  return 0;
}
