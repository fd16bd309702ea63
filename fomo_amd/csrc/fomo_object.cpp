// fomo_object.cpp — the FoMo C++ API (include/FoMo.h) on top of the MI355X engine (fomo_hip.h).
//
// Host containers keep the reference's value semantics. FoMoObject::render() uploads the datacube
// once (again only after push_back_datapoint/setdata), computes the emission on the GPU, renders
// the (l,b) views on the GPU (the NearestNeighbour views of one call together, in one captured
// batch) and builds the RenderCube grids on the host exactly as fomo-nearestneighbour.cpp:233-279
// / fomo-thomson.cpp:348-364 do, while a background thread writes the previous view's files. The
// emission cube stays on the device: the GoftCube is read back only when readgoftcube() asks for it
// (or before the datacube changes). Error behaviour follows the reference facade: missing tables,
// a wrong instrument table and unknown methods print to stderr and exit(EXIT_FAILURE); the C ABI
// below returns codes instead.

#include <sys/stat.h>

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>

#include "../../include/FoMo.h"
#include "../../include/fomo_hip.h"
#include "fomo_writer.h"

namespace {

const double kC = 2.99792458e8;  // GSL_CONST_MKSA_SPEED_OF_LIGHT
const double kPi = M_PI;
const double kMmPerArcsec = 0.715;

[[noreturn]] void die(const std::string& msg) {
  std::cerr << msg << std::endl << std::flush;
  std::exit(EXIT_FAILURE);
}

void engine() {
  static bool ok = false;
  if (ok) return;
  int rc = fomo_init(0);
  if (rc != FOMO_OK) die(std::string("Error: GPU engine initialisation failed: ") + fomo_last_error());
  ok = true;
}

// A dataset or emission cube on the device, destroyed with its last owner.
struct DeviceDataset {
  int id = -1;
  ~DeviceDataset() {
    if (id >= 0) fomo_dataset_destroy(id);
  }
};
struct DeviceEmission {
  int id = -1;
  std::shared_ptr<DeviceDataset> ds;  // released after the emission cube
  ~DeviceEmission() {
    if (id >= 0) fomo_goft_destroy(id);
  }
};

}  // namespace

struct FoMo::FoMoDeviceState {
  std::shared_ptr<DeviceDataset> dataset;    // upload of the datacube; null when none or stale
  std::shared_ptr<DeviceEmission> emission;  // emission cube of the last render()
  bool goft_pending = false;                 // goftcube holds only the metadata of `emission`
  bool thomson = false;
  std::vector<std::string> goft_units;
};

namespace {

// The state of one object alone: copies of a FoMoObject share it until one of them changes it.
FoMo::FoMoDeviceState& own_state(std::shared_ptr<FoMo::FoMoDeviceState>& d) {
  if (!d) d = std::make_shared<FoMo::FoMoDeviceState>();
  else if (d.use_count() > 1) d = std::make_shared<FoMo::FoMoDeviceState>(*d);
  return *d;
}

// Fill the GoftCube of the last render() from the device (emissionfromdatacube's output,
// fomo-CHIANTI.cpp:296-321, or thomsonfromdatacube's, fomo-thomson.cpp:60-92): the datacube's
// grid, the two emission variables in the original point order, and for a spectral emission the
// velocity variables.
void materialize(FoMo::FoMoDeviceState& st, const FoMo::DataCube& dc, FoMo::GoftCube& gc) {
  if (!st.goft_pending) return;
  st.goft_pending = false;
  const int64_t n = dc.readngrid();
  std::vector<float> a(n), c(n);
  if (fomo_goft_read(st.emission->id, a.data(), c.data()) != FOMO_OK)
    die(std::string("Error: emission read-back failed: ") + fomo_last_error());
  FoMo::tgrid g = dc.readgrid();
  FoMo::tvars v{std::move(a), std::move(c)};
  if (st.thomson) {
    gc.setdata(g, v, nullptr);
    return;
  }
  for (int i = 2; i < dc.readnvars(); ++i) v.push_back(dc.readvar(i));
  std::vector<std::string> units = st.goft_units;
  gc.setdata(g, v, &units);
}

struct TableEntry {
  int id = -1;
  std::string ion, unit;
  double lambda0 = 0, atweight = 0;
  long mtime = 0;
};
std::map<std::string, TableEntry> g_tables;

long file_mtime(const std::string& p) {
  struct stat st;
  if (stat(p.c_str(), &st) != 0) return -1;
  return (long)st.st_mtime;
}

// The G(T) table at path, parsed and uploaded once (again when the file changes). An imaging
// render needs an instrument table, whose ion field is the channel (fomo-CHIANTI.cpp:278-282);
// every failure here ends the program before the GPU is initialised.
const TableEntry& load_table(const std::string& path, bool imaging) {
  const auto check_ion = [&](const std::string& ion, double l0) {
    if (imaging && std::atoi(ion.c_str()) != int(l0 + 0.5)) die("GOFT table is not correct!");
  };
  long mt = file_mtime(path);
  if (mt < 0) die("Error: no CHIANTI G(T) file exists: " + path);
  auto it = g_tables.find(path);
  if (it != g_tables.end() && it->second.mtime == mt) {
    check_ion(it->second.ion, it->second.lambda0);
    return it->second;
  }
  char ion[256], unit[1024];
  double l0 = 0, aw = 0;
  int nt = 0, nr = 0;
  if (fomo_table_read(path.c_str(), ion, sizeof ion, &l0, &aw, unit, sizeof unit, &nt, &nr, nullptr, nullptr,
                      nullptr) != FOMO_OK)
    die("Error: CHIANTI table not read in properly!");
  std::vector<float> lt(nt), ln(nr), g((size_t)nt * nr);
  if (fomo_table_read(path.c_str(), ion, sizeof ion, &l0, &aw, unit, sizeof unit, &nt, &nr, lt.data(), ln.data(),
                      g.data()) != FOMO_OK)
    die("Error: CHIANTI table not read in properly!");
  check_ion(ion, l0);
  engine();
  TableEntry e;
  if (it != g_tables.end() && it->second.id >= 0) fomo_table_destroy(it->second.id);
  if (fomo_table_create(nt, nr, lt.data(), ln.data(), g.data(), &e.id) != FOMO_OK)
    die(std::string("Error: table upload failed: ") + fomo_last_error());
  e.ion = ion;
  e.unit = unit;
  e.lambda0 = l0;
  e.atweight = aw;
  e.mtime = mt;
  g_tables[path] = e;
  return g_tables[path];
}

// fomo-CHIANTI.cpp:23-62: "Z value element" triples ended by a negative Z; 10^(ion - H).
double abund_from_file(const std::string& path, const std::string& ion) {
  std::ifstream in(path);
  if (!in) die("Error: no CHIANTI abundance file exists\n");
  const std::string ionname = ion.substr(0, ion.find_first_of('_'));
  double hvalue = 0, ionvalue = 0, value;
  int nelement = -1;
  std::string element;
  if (!(in >> nelement)) nelement = -1;
  while (nelement >= 0) {
    if (!(in >> value >> element)) break;
    if (element.size() == 1 && strncasecmp(element.c_str(), "h", 1) == 0) hvalue = value;
    if (element.size() == ionname.size() && strncasecmp(element.c_str(), ionname.c_str(), ionname.size()) == 0)
      ionvalue = value;
    if (!(in >> nelement)) break;
  }
  if (hvalue == 0. || ionvalue == 0.) die("Error: invalid CHIANTI abundance file\n");
  return std::pow(10., ionvalue - hvalue);
}

// The tables were computed with sun_coronal_2012_schmelz.abund. The reference embeds it as a
// string; here it is read from $FOMO_DEFAULT_ABUND or from beside the G(T) table.
std::string default_abund_path(const std::string& chiantifile) {
  const char* env = std::getenv("FOMO_DEFAULT_ABUND");
  if (env && *env) return env;
  std::string dir = ".";
  size_t p = chiantifile.rfind('/');
  if (p != std::string::npos) dir = chiantifile.substr(0, p);
  return dir + "/sun_coronal_2012_schmelz.abund";
}

std::string angle_tag(double a) {
  std::stringstream ss;
  ss << std::setfill('0') << std::setw(3) << std::round(a / kPi * 180.);
  return ss.str();
}

}  // namespace

// =====================================================================================
// element-wise operations (fomo-operations.cpp:8-114)
// =====================================================================================
FoMo::tphysvar FoMo::pow(const double base, FoMo::tphysvar const& in) {
  FoMo::tphysvar out(in.size());
  for (size_t i = 0; i < in.size(); ++i) out[i] = std::pow(base, in[i]);
  return out;
}
FoMo::tphysvar FoMo::operator/(FoMo::tphysvar const& a, FoMo::tphysvar const& b) {
  assert(a.size() == b.size());
  FoMo::tphysvar out(a.size());
  for (size_t i = 0; i < a.size(); ++i) out[i] = a[i] / b[i];
  return out;
}
FoMo::tphysvar FoMo::operator*(FoMo::tphysvar const& a, FoMo::tphysvar const& b) {
  assert(a.size() == b.size());
  FoMo::tphysvar out(a.size());
  for (size_t i = 0; i < a.size(); ++i) out[i] = a[i] * b[i];
  return out;
}
FoMo::tphysvar FoMo::operator*=(FoMo::tphysvar const& a, FoMo::tphysvar const& b) { return a * b; }
FoMo::tphysvar FoMo::log10(FoMo::tphysvar const& in) {
  FoMo::tphysvar out(in.size());
  for (size_t i = 0; i < in.size(); ++i) {
    if (in[i] <= 0) {
      std::cerr << "Warning: some densities were <= 0 at position " << i << std::endl;
      out[i] = 0;
    } else {
      out[i] = std::log10(in[i]);
    }
  }
  return out;
}
FoMo::tphysvar FoMo::sqrt(FoMo::tphysvar const& in) {
  FoMo::tphysvar out(in.size());
  for (size_t i = 0; i < in.size(); ++i) out[i] = std::sqrt(in[i]);
  return out;
}
FoMo::tphysvar FoMo::operator*(double const& c, FoMo::tphysvar const& in) {
  FoMo::tphysvar out(in.size());
  for (size_t i = 0; i < in.size(); ++i) out[i] = c * in[i];
  return out;
}

// =====================================================================================
// DataCube (fomo-datacube.cpp)
// =====================================================================================
FoMo::DataCube::DataCube(const int indim) : dim(indim), nvars(0), ng(0) {
  grid.resize(indim);
  unit.resize(indim);
}
FoMo::DataCube::~DataCube() {}
int FoMo::DataCube::readdim() const { return dim; }
int FoMo::DataCube::readngrid() const { return ng; }
int FoMo::DataCube::readnvars() const { return nvars; }
std::vector<std::string> FoMo::DataCube::readunit() const { return unit; }
FoMo::tgrid FoMo::DataCube::readgrid() const { return grid; }
FoMo::tphysvar FoMo::DataCube::readvar(const unsigned int nvar) const {
  assert(nvar < nvars);
  return vars.at(nvar);
}
void FoMo::DataCube::setdim(const int indim) {
  dim = indim;
  grid.resize(indim);
  unit.resize(indim + nvars);
}
void FoMo::DataCube::setnvars(const int innvars) {
  assert(innvars >= 0);
  nvars = innvars;
  vars.resize(innvars);
  unit.resize(dim + nvars);
}
void FoMo::DataCube::setngrid(const int inngrid) {
  assert(inngrid >= 0);
  ng = inngrid;
  for (unsigned int i = 0; i < dim; ++i) grid.at(i).resize(inngrid);
}
void FoMo::DataCube::setgrid(FoMo::tgrid ingrid, std::vector<std::string>* inunit) {
  assert(!ingrid.empty());
  std::vector<std::string> oldunits;
  if (unit.size() > dim) oldunits.assign(unit.begin() + dim, unit.end());
  dim = ingrid.size();
  const unsigned int tempng = ingrid[0].size();
  for (unsigned int i = 1; i < dim; ++i) assert(ingrid[i].size() == tempng);
  grid = std::move(ingrid);
  ng = tempng;
  if (inunit) {
    assert(inunit->size() == dim);
    unit = *inunit;
    unit.insert(unit.end(), oldunits.begin(), oldunits.end());
  } else if (unit.size() < dim + nvars) {
    unit.resize(dim + nvars);
  }
}
void FoMo::DataCube::setvar(const unsigned int nvar, const FoMo::tphysvar var, std::string inunit) {
  assert(nvar < vars.size());
  vars.at(nvar) = var;
  if (unit.size() < dim + nvars) unit.resize(dim + nvars);
  unit.at(dim + nvar) = inunit;
}
void FoMo::DataCube::push_back(std::vector<double> coordinate, std::vector<double> variables,
                               std::vector<std::string>* unitvec) {
  if (ng == 0) {
    tgrid g;
    for (double c : coordinate) g.push_back(tcoord{(float)c});
    tvars v;
    for (double x : variables) v.push_back(tphysvar{(float)x});
    setdata(g, v, unitvec);
  } else {
    assert(coordinate.size() == dim);
    assert(variables.size() == nvars);
    for (unsigned int i = 0; i < dim; ++i) grid.at(i).push_back((float)coordinate.at(i));
    for (unsigned int i = 0; i < nvars; ++i) vars.at(i).push_back((float)variables.at(i));
    ng++;
  }
}
void FoMo::DataCube::setdata(tgrid& ingrid, tvars& indata, std::vector<std::string>* unitvec) {
  setgrid(ingrid);
  if (!indata.empty()) {
    setnvars(indata.size());
    for (unsigned int i = 0; i < nvars; ++i) {
      assert(indata[i].size() == ng);
      vars.at(i) = indata[i];
    }
  }
  if (unitvec) {
    assert(unitvec->size() == dim + nvars);
    unit = *unitvec;
  }
  if (unit.size() < dim + nvars) unit.resize(dim + nvars);
}

// =====================================================================================
// GoftCube (fomo-goftcube.cpp; writer in fomo_io.cpp)
// =====================================================================================
FoMo::GoftCube::GoftCube(const int indim) : DataCube(indim) {
  chiantifile = "../chiantitables/goft_table_fe_12_0194_abco.dat";
  abundfile = "/empty";
  lambda0 = 193.509;
  writeoptions = std::bitset<noptions>(std::string("0010"));
}
FoMo::GoftCube::GoftCube(FoMo::DataCube incube) : DataCube(incube.readdim()) {
  ng = incube.readngrid();
  grid = incube.readgrid();
  nvars = 0;
  chiantifile = "../chiantitables/goft_table_fe_12_0194_abco.dat";
  abundfile = "/empty";
  lambda0 = 193.509;
  writeoptions = std::bitset<noptions>(std::string("0010"));
}
std::string FoMo::GoftCube::readabundfile() { return abundfile; }
std::string FoMo::GoftCube::readchiantifile() { return chiantifile; }
void FoMo::GoftCube::setabundfile(const std::string s) { abundfile = s; }
void FoMo::GoftCube::setchiantifile(const std::string s) { chiantifile = s; }
void FoMo::GoftCube::setlambda0(const double l) { lambda0 = l; }
double FoMo::GoftCube::readlambda0() { return lambda0; }
std::bitset<FoMo::noptions> FoMo::GoftCube::getwriteoptions() { return writeoptions; }
void FoMo::GoftCube::setwriteoptions(std::bitset<FoMo::noptions> o) { writeoptions = o; }
void FoMo::GoftCube::setwriteoutbinary(const bool v) { writeoptions[0] = v; }
void FoMo::GoftCube::setwriteouttext(const bool v) { writeoptions[1] = v; }
void FoMo::GoftCube::setwriteoutzip(const bool v) { writeoptions[2] = v; }
void FoMo::GoftCube::setwriteoutdeletefiles(const bool v) { writeoptions[3] = v; }

// =====================================================================================
// RenderCube (fomo-rendercube.cpp)
// =====================================================================================
FoMo::RenderCube::RenderCube(FoMo::GoftCube goftcube) : GoftCube(goftcube.readdim()) {
  chiantifile = goftcube.readchiantifile();
  abundfile = goftcube.readabundfile();
  lambda0 = goftcube.readlambda0();
  grid = goftcube.readgrid();
  ng = goftcube.readngrid();
  writeoptions = goftcube.getwriteoptions();
  l = 0;
  b = 0;
  x_pixel = 101;
  y_pixel = 102;
  z_pixel = 300;
  lambda_pixel = 30;
  lambda_width = 200000;
  rendermethod = "NearestNeighbour";
  observationtype = Spectroscopic;
}
void FoMo::RenderCube::readresolution(int& nx, int& ny, int& nz, int& nl, double& w) {
  nx = x_pixel;
  ny = y_pixel;
  nz = z_pixel;
  nl = lambda_pixel;
  w = lambda_width;
}
void FoMo::RenderCube::setresolution(const int& nx, const int& ny, const int& nz, const int& nl, const double& w) {
  x_pixel = nx;
  y_pixel = ny;
  z_pixel = nz;
  lambda_pixel = nl;
  lambda_width = w;
  // fomo-rendercube.cpp:74-84: widths below 1 are taken as Angstrom and converted with lambda0
  if (w < 1.) {
    std::cout << "******Warning! In newer versions of FoMo, the spectral window is given in m/s, rather than Angstrom."
              << std::endl;
    lambda_width = w * kC / lambda0;
    std::cout << "******The input value of " << w << " was converted to " << lambda_width << "m/s" << std::endl
              << std::flush;
  } else if (w < 1000.) {
    std::cout << "******Warning! Your spectral window width was set to a small value. Are you sure it is in m/s?"
              << std::endl
              << std::flush;
  }
}
void FoMo::RenderCube::setrendermethod(const std::string m) { rendermethod = m; }
std::string FoMo::RenderCube::readrendermethod() { return rendermethod; }
void FoMo::RenderCube::setobservationtype(FoMo::FoMoObservationType t) { observationtype = t; }
FoMo::FoMoObservationType FoMo::RenderCube::readobservationtype() { return observationtype; }
void FoMo::RenderCube::setangles(const double lin, const double bin) {
  l = lin;
  b = bin;
}
void FoMo::RenderCube::readangles(double& lo, double& bo) {
  lo = l;
  bo = b;
}

// =====================================================================================
// FoMoObject (fomo-object.cpp)
// =====================================================================================
FoMo::FoMoObject::FoMoObject(const int indim) : datacube(indim), goftcube(datacube), rendering(goftcube) {}

FoMo::FoMoObject::~FoMoObject() {}

void FoMo::FoMoObject::setrendermethod(const std::string m) { rendering.setrendermethod(m); }
std::string FoMo::FoMoObject::readrendermethod() { return rendering.readrendermethod(); }
FoMo::DataCube FoMo::FoMoObject::readdatacube() { return datacube; }
FoMo::GoftCube FoMo::FoMoObject::readgoftcube() {
  if (device && device->goft_pending) materialize(own_state(device), datacube, goftcube);
  return goftcube;
}
FoMo::RenderCube FoMo::FoMoObject::readrendering() { return rendering; }
void FoMo::FoMoObject::setabundfile(const std::string s) {
  goftcube.setabundfile(s);
  rendering.setabundfile(s);
}
std::string FoMo::FoMoObject::readabundfile() { return goftcube.readabundfile(); }

// fomo-object.cpp:110-116 with readgoftfromchianti(file) = first two tokens (fomo-CHIANTI.cpp:123-141)
void FoMo::FoMoObject::setchiantifile(const std::string inchianti) {
  std::ifstream in(inchianti);
  if (!in) die("Error: no CHIANTI G(T) file exists: " + inchianti + "\n");
  std::string ion;
  double l0 = 0;
  in >> ion >> l0;
  goftcube.setlambda0(l0);
  goftcube.setchiantifile(inchianti);
  rendering.setlambda0(l0);
  rendering.setchiantifile(inchianti);
}
std::string FoMo::FoMoObject::readchiantifile() { return goftcube.readchiantifile(); }

void FoMo::FoMoObject::push_back_datapoint(std::vector<double> coordinate, std::vector<double> variables,
                                           std::vector<std::string>* unitvec) {
  // the reference ignores unitvec here (fomo-object.cpp:139-141)
  (void)unitvec;
  FoMoDeviceState& st = own_state(device);
  materialize(st, datacube, goftcube);  // the last render's GoftCube keeps the data it was made from
  datacube.push_back(coordinate, variables);
  st.dataset.reset();
}

void FoMo::FoMoObject::setdata(tgrid& ingrid, tvars& indata, std::vector<std::string>* unitvec) {
  FoMoDeviceState& st = own_state(device);
  materialize(st, datacube, goftcube);
  datacube.setdata(ingrid, indata, unitvec);
  st.dataset.reset();
}

void FoMo::FoMoObject::setobservationtype(FoMoObservationType t) { rendering.setobservationtype(t); }
FoMo::FoMoObservationType FoMo::FoMoObject::readobservationtype() { return rendering.readobservationtype(); }
void FoMo::FoMoObject::setresolution(const int& nx, const int& ny, const int& nz, const int& nl, const double& w) {
  rendering.setresolution(nx, ny, nz, nl, w);
}
void FoMo::FoMoObject::readresolution(int& nx, int& ny, int& nz, int& nl, double& w) {
  rendering.readresolution(nx, ny, nz, nl, w);
}
void FoMo::FoMoObject::setrenderingdata(tgrid ingrid, tvars invars) { rendering.setdata(ingrid, invars); }
void FoMo::FoMoObject::setoutfile(const std::string s) { outfile = s; }
void FoMo::FoMoObject::setwriteoptions(std::bitset<noptions> o) {
  rendering.setwriteoptions(o);
  goftcube.setwriteoptions(o);
}
void FoMo::FoMoObject::setwriteoutbinary(const bool v) {
  rendering.setwriteoutbinary(v);
  goftcube.setwriteoutbinary(v);
}
void FoMo::FoMoObject::setwriteouttext(const bool v) {
  rendering.setwriteouttext(v);
  goftcube.setwriteouttext(v);
}
void FoMo::FoMoObject::setwriteoutzip(const bool v) {
  rendering.setwriteoutzip(v);
  goftcube.setwriteoutzip(v);
}
void FoMo::FoMoObject::setwriteoutdeletefiles(const bool v) {
  rendering.setwriteoutdeletefiles(v);
  goftcube.setwriteoutdeletefiles(v);
}

namespace {

std::shared_ptr<DeviceDataset> upload_dataset(const FoMo::DataCube& dc) {
  engine();
  const int dim = dc.readdim();
  const int nv = dc.readnvars();
  const int64_t n = dc.readngrid();
  if (n < 1) die("Error: the datacube holds no points.");
  FoMo::tgrid g = dc.readgrid();
  std::vector<FoMo::tphysvar> vars(nv);
  std::vector<const float*> vp(nv);
  for (int i = 0; i < nv; ++i) {
    vars[i] = dc.readvar(i);
    vp[i] = vars[i].data();
  }
  auto d = std::make_shared<DeviceDataset>();
  int rc = fomo_dataset_create(n, dim == 2 ? 2 : 3, g[0].data(), g[1].data(), dim >= 3 ? g[2].data() : nullptr, nv,
                               vp.data(), 0.0, &d->id);
  if (rc != FOMO_OK) die(std::string("Error: dataset upload failed: ") + fomo_last_error());
  return d;
}

// Views of one NearestNeighbour render() batched together: at most this much image per batch.
constexpr double kBatchImageBytes = 2.0 * (1 << 30);

}  // namespace

// fomo-object.cpp:374-440
void FoMo::FoMoObject::render(const std::vector<double> lvec, const std::vector<double> bvec) {
  int nx, ny, nz, nl;
  double width;
  rendering.readresolution(nx, ny, nz, nl, width);
  rendering.setobservationtype(nl <= 1 ? Imaging : Spectroscopic);
  const std::string method = rendering.readrendermethod();
  const bool thomson = method == "Thomson";
  // "CGAL": the tetra path (fomo-CGAL.cpp:309-338); "Linear": its barycentric variant (extension)
  const bool tet = method == "CGAL" || method == "Linear";
  const bool proj = method == "Projection";  // fomo-projection.cpp:218-239
  if (method != "NearestNeighbour" && !thomson && !tet && !proj) {
    if (method == "CGAL2D") die("Error: render method " + method + " is not available in this build.");
    die("Error: unknown rendering method.");
  }
  const std::bitset<noptions> woptions = goftcube.getwriteoptions();
  // the G(T) table first, as emissionfromdatacube reads it before anything else
  // (fomo-CHIANTI.cpp:244): a missing or malformed table, or an imaging table of another
  // instrument (:278-282), ends the program before the GPU is used
  const TableEntry* te = nullptr;
  const bool imaging = rendering.readobservationtype() == Imaging;
  if (!thomson) te = &load_table(rendering.readchiantifile(), imaging);
  FoMoDeviceState& st = own_state(device);
  if (!st.dataset) st.dataset = upload_dataset(datacube);
  const int ds = st.dataset->id;
  const int dim = datacube.readdim();

  // ---- emission (the GoftCube): stays on the device, read back by readgoftcube() ------------
  auto em = std::make_shared<DeviceEmission>();
  em->ds = st.dataset;
  FoMo::GoftCube meta(dim);  // the GoftCube's metadata; its data arrive on demand
  std::string emunit;
  std::vector<std::string> units;
  if (thomson) {
    if (fomo_emission_thomson(ds, &em->id) != FOMO_OK) die(std::string("Error: ") + fomo_last_error());
    meta.setchiantifile("thomson scattering");
    meta.setabundfile("thomson scattering");
    meta.setlambda0(0.);
  } else {
    const std::string cfile = rendering.readchiantifile();
    const TableEntry& tab = *te;
    double abundratio = 1.0;
    const std::string afile = rendering.readabundfile();
    if (!imaging && afile != std::string("/empty"))
      abundratio = abund_from_file(afile, tab.ion) / abund_from_file(default_abund_path(cfile), tab.ion);
    if (datacube.readnvars() < 2) die("Error: the datacube needs at least n and T.");
    if (fomo_emission_chianti(ds, tab.id, tab.lambda0, tab.atweight, imaging ? 1 : 0, abundratio, &em->id) != FOMO_OK)
      die(std::string("Error: ") + fomo_last_error());
    std::vector<std::string> dunits = datacube.readunit();
    for (int i = 0; i < dim; ++i) units.push_back(i < (int)dunits.size() ? dunits[i] : "");
    units.push_back(tab.unit);
    units.push_back("\\AA{}");
    // the velocity columns take dataunits.at(i) as the reference indexes them (fomo-CHIANTI.cpp:314):
    // in 3D the z grid unit and the first two variables' units; empty where .at() would throw
    for (int i = 2; i < datacube.readnvars(); ++i) units.push_back(i < (int)dunits.size() ? dunits[i] : "");
    meta.setchiantifile(cfile);
    meta.setabundfile(afile);
    meta.setlambda0(tab.lambda0);
    emunit = tab.unit;
  }
  meta.setwriteoptions(woptions);
  st.emission = em;
  st.goft_pending = true;
  st.thomson = thomson;
  st.goft_units = units;
  goftcube = meta;
  const int goft = em->id;
  const double lambda0 = goftcube.readlambda0();

  const bool dn = emunit.find("DN") != std::string::npos;
  const int nlo = std::max(nl, 1);
  const size_t npix = (size_t)nx * ny * nlo;
  std::vector<std::pair<double, double>> views;
  for (double lv : lvec)
    for (double bv : bvec) views.emplace_back(lv, bv);
  if (views.empty()) {
    // no view: the reference keeps RenderCube(goftcube) (fomo-nearestneighbour.cpp:287)
    materialize(st, datacube, goftcube);
    FoMo::RenderCube r(goftcube);
    r.setrendermethod(method);
    r.setobservationtype(rendering.readobservationtype());
    rendering = r;
    return;
  }

  // One view's RenderCube, built on the host from its image and meta (extents, scaling).
  const auto make_cube = [&](const std::vector<float>& img, const double* m, double lv, double bv) {
    const double minx = m[0], maxx = m[1], miny = m[2], maxy = m[3];
    FoMo::RenderCube rc3(meta);
    FoMo::tgrid newgrid;
    FoMo::tvars newdata;
    if (!thomson) {
      // fomo-nearestneighbour.cpp:124-128, 152-153, 200-211, 233-279 (the CGAL path builds the same
      // grids, fomo-CGAL.cpp:187-188, 210-218, 262-285, and so does the projection,
      // fomo-projection.cpp:76-118, 161-212, whose view cubes also say "NearestNeighbour", :203)
      FoMo::tcoord xv(npix), yv(npix), lv3(nl > 1 ? npix : 0);
      const double wla = width * lambda0 / kC;
      for (int i = 0; i < ny; ++i) {
        const double y = double(i) / (ny - 1) * (maxy - miny) + miny;
        for (int j = 0; j < nx; ++j) {
          const double x = double(j) / (nx - 1) * (maxx - minx) + minx;
          for (int il = 0; il < nlo; ++il) {
            const size_t ind = ((size_t)i * nx + j) * nlo + il;
            xv[ind] = (float)x;
            yv[ind] = (float)y;
            if (nl > 1) lv3[ind] = (float)(double(il) / (nl - 1) * wla - wla / 2. + lambda0);
          }
        }
      }
      std::vector<std::string> vunits{"Mm", "Mm"};
      if (nl > 1) vunits.push_back("\\AA{}");
      vunits.push_back("erg cm^{-2} s^{-1} \\AA{}^{-1}");
      if (dn) {
        vunits[0] = "arcsec";
        xv = FoMo::operator*(1. / kMmPerArcsec, xv);
        vunits[1] = "arcsec";
        yv = FoMo::operator*(1. / kMmPerArcsec, yv);
        vunits.back() = "DN s^{-1} pixel^{-1}";
      }
      newgrid.push_back(std::move(xv));
      newgrid.push_back(std::move(yv));
      if (nl > 1) newgrid.push_back(std::move(lv3));
      newdata.push_back(img);
      rc3.setdata(newgrid, newdata, &vunits);
      rc3.setrendermethod(tet ? method : std::string("NearestNeighbour"));
      rc3.setresolution(nx, ny, nz, nl, width);
      rc3.setobservationtype(nl == 1 ? FoMo::Imaging : FoMo::Spectroscopic);
    } else {
      // fomo-thomson.cpp:228-232, 333-334, 348-364
      FoMo::tcoord xv((size_t)nx * ny), yv((size_t)nx * ny);
      for (int i = 0; i < ny; ++i) {
        const double y = double(i) / (ny - 1) * (maxy - miny) + miny;
        for (int j = 0; j < nx; ++j) {
          xv[(size_t)i * nx + j] = (float)(double(j) / (nx - 1) * (maxx - minx) + minx);
          yv[(size_t)i * nx + j] = (float)y;
        }
      }
      newgrid.push_back(std::move(xv));
      newgrid.push_back(std::move(yv));
      newdata.push_back(std::vector<float>(img.begin(), img.begin() + (size_t)nx * ny));
      std::vector<std::string> tunits{"Mm", "Mm", "pB"};
      rc3.setdata(newgrid, newdata, &tunits);
      rc3.setrendermethod("Thomson");
      rc3.setresolution(nx, ny, nz, 1, 100000.);
      rc3.setobservationtype(FoMo::Imaging);
    }
    rc3.setangles(lv, bv);
    return rc3;
  };

  // Per-view files (fomo-nearestneighbour.cpp:294-302: one per view, even when outfile is "") are
  // written by a background thread while the next view is rendered and built.
  FoMo::RenderCube last(meta);
  {
    fomo::CubeWriter writer(2);
    const auto emit = [&](const std::vector<float>& img, const double* m, size_t v) {
      const double lv = views[v].first, bv = views[v].second;
      FoMo::RenderCube rc3 = make_cube(img, m, lv, bv);
      if (v + 1 == views.size()) last = rc3;
      writer.submit(std::move(rc3), outfile + "l" + angle_tag(lv) + "b" + angle_tag(bv) + ".txt");
    };
    std::vector<float> img(npix);
    double m[10];
    const bool batched = !thomson && !tet && !proj && views.size() > 1;
    if (batched) {
      // all NearestNeighbour views in captured batches of at most kBatchImageBytes of images;
      // a batch gives each view the image fomo_render_nn gives it
      const size_t per = std::max<size_t>(1, (size_t)(kBatchImageBytes / (4.0 * npix)));
      for (size_t v0 = 0; v0 < views.size(); v0 += per) {
        const size_t nv = std::min(per, views.size() - v0);
        std::vector<double> ls(nv), bs(nv);
        for (size_t k = 0; k < nv; ++k) ls[k] = views[v0 + k].first, bs[k] = views[v0 + k].second;
        int bt = -1;
        if (fomo_batch_create(1, &goft, (int)nv, ls.data(), bs.data(), nx, ny, nz, nl, width, dn ? 1 : 0, &bt) !=
                FOMO_OK ||
            fomo_batch_run(bt) != FOMO_OK)
          die(std::string("Error: render failed: ") + fomo_last_error());
        for (size_t k = 0; k < nv; ++k) {
          if (fomo_batch_read_view(bt, (int)k, 0, img.data(), m) != FOMO_OK)
            die(std::string("Error: render failed: ") + fomo_last_error());
          emit(img, m, v0 + k);
        }
        fomo_batch_destroy(bt);
      }
    } else {
      int64_t tet0[2] = {0, 0};
      if (tet) fomo_dataset_tet_stats(ds, tet0);
      for (size_t v = 0; v < views.size(); ++v) {
        const double lv = views[v].first, bv = views[v].second;
        int rc = thomson ? fomo_render_thomson(goft, lv, bv, nx, ny, nz, img.data(), m)
                 : tet   ? fomo_render_tet(goft, lv, bv, nx, ny, nz, nl, width, dn ? 1 : 0,
                                           method == "Linear" ? 1 : 0, img.data(), m)
                 : proj  ? fomo_render_projection(goft, lv, bv, nx, ny, nz, nl, width, dn ? 1 : 0, img.data(), m)
                         : fomo_render_nn(goft, lv, bv, nx, ny, nz, nl, width, dn ? 1 : 0, img.data(), m);
        if (rc != FOMO_OK) die(std::string("Error: render failed: ") + fomo_last_error());
        emit(img, m, v);
      }
      int64_t tet1[2] = {0, 0};
      if (tet && fomo_dataset_tet_stats(ds, tet1) == FOMO_OK && tet1[1] > tet0[1])
        std::cerr << "Warning: " << (tet1[1] - tet0[1])
                  << " LOS samples could not be located exactly in the triangulation" << std::endl;
    }
    writer.drain();
  }
  last.setrendermethod(method);
  last.setobservationtype(rendering.readobservationtype());
  rendering = last;
}

void FoMo::FoMoObject::render(const double l, const double b) {
  std::vector<double> lvec{l};
  std::vector<double> bvec{b};
  render(lvec, bvec);
}
