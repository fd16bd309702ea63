// FoMo.h — public C++ API of the FoMo render library, MI355X build.
//
// Drop-in for the reference header (fomo-c/src/FoMo.h:15-219): same namespace, types, classes,
// member signatures and result layout, so programs written against FoMo-C (for example
// fomo-c/example/example.cpp) compile and link unchanged against libFoMo.so. Rendering runs on
// the GPU through the C ABI in fomo_hip.h; the containers below stay host-side value types.
#ifndef FOMO_H
#define FOMO_H

#include <bitset>
#include <memory>
#include <string>
#include <vector>

namespace FoMo {

// Coordinates and variables are stored in single precision, as in the reference (FoMo.h:20-32).
typedef std::vector<float> tcoord;
typedef std::vector<tcoord> tgrid;
typedef std::vector<float> tphysvar;
typedef std::vector<tphysvar> tvars;

// Element-wise helpers (fomo-operations.cpp:8-114).
tphysvar pow(const double, tphysvar const&);
tphysvar operator/(tphysvar const&, tphysvar const&);
tphysvar operator*(tphysvar const&, tphysvar const&);
tphysvar operator*=(tphysvar const&, tphysvar const&);
tphysvar log10(tphysvar const&);
tphysvar operator*(double const&, tphysvar const&);
tphysvar sqrt(tphysvar const&);

// Simulation data: dim coordinates and nvars variables per point (FoMo.h:48-76).
class DataCube {
 protected:
  unsigned int dim;
  unsigned int nvars;
  unsigned int ng;
  FoMo::tgrid grid;
  FoMo::tvars vars;
  std::vector<std::string> unit;
  void setgrid(tgrid ingrid, std::vector<std::string>* inunit = NULL);
  void setvar(const unsigned int, const tphysvar, std::string inunit = "");

 public:
  DataCube(const int = 3);
  ~DataCube();
  int readdim() const;
  int readngrid() const;
  int readnvars() const;
  std::vector<std::string> readunit() const;
  tgrid readgrid() const;
  tphysvar readvar(const unsigned int) const;
  void setdim(const int indim);
  void setnvars(const int innvars);
  void setngrid(const int inngrid);
  void setdata(tgrid& ingrid, tvars& indata, std::vector<std::string>* unitvec = NULL);
  void push_back(std::vector<double> coordinate, std::vector<double> variables,
                 std::vector<std::string>* unitvec = NULL);
};

const int noptions = 4;  // write options: binary, text, zip, delete unzipped

// Emission data ready for rendering (FoMo.h:86-113).
class GoftCube : public DataCube {
 protected:
  std::string chiantifile;
  std::string abundfile;
  double lambda0;
  std::bitset<noptions> writeoptions;

 public:
  GoftCube(const int = 3);
  GoftCube(DataCube datacube);
  void setchiantifile(const std::string inchianti);
  void setabundfile(const std::string inabund);
  std::string readchiantifile();
  std::string readabundfile();
  double readlambda0();
  void setlambda0(const double lambda0);
  void writegoftcube(const std::string);
  void readgoftcube(const std::string);
  std::bitset<noptions> getwriteoptions();
  void setwriteoptions(std::bitset<noptions> options);
  void setwriteoutbinary(const bool = true);
  void setwriteouttext(const bool = true);
  void setwriteoutzip(const bool = true);
  void setwriteoutdeletefiles(const bool = true);
};

enum FoMoObservationType { ObservationTypeNotDefined, Spectroscopic, Imaging };

// Rendering result: grid (x, y[, lambda]) and intensity, index (iy*nx+jx)*nl+il (FoMo.h:134-156).
class RenderCube : public GoftCube {
 protected:
  double l;
  double b;
  int x_pixel;
  int y_pixel;
  int z_pixel;
  int lambda_pixel;
  double lambda_width;
  std::string rendermethod;
  FoMoObservationType observationtype;

 public:
  RenderCube(GoftCube goftcube);
  void setresolution(const int& x_pixel, const int& y_pixel, const int& z_pixel, const int& lambda_pixel,
                     const double& lambda_width);
  void readresolution(int& x_pixel, int& y_pixel, int& z_pixel, int& lambda_pixel, double& lambda_width);
  void setangles(const double l, const double b);
  void readangles(double& l, double& b);
  void setrendermethod(const std::string inrendermethod);
  std::string readrendermethod();
  void setobservationtype(FoMoObservationType);
  FoMoObservationType readobservationtype();
};

// Device-side state of a FoMoObject (defined in the library).
struct FoMoDeviceState;

// Main object: datacube -> goftcube -> rendering (FoMo.h:166-218).
class FoMoObject {
 protected:
  std::string outfile;
  FoMo::DataCube datacube;
  FoMo::GoftCube goftcube;
  FoMo::RenderCube rendering;

 private:
  // Not in the reference: the datacube's upload to the GPU and the emission cube of the last
  // render(), shared by copies of the object until one of them changes its data.
  std::shared_ptr<FoMoDeviceState> device;

 public:
  FoMoObject(const int = 3);
  ~FoMoObject();
  void render(const double = 0, const double = 0);
  void render(const std::vector<double> lvec, const std::vector<double> bvec);
  void setrenderingdata(tgrid ingrid, tvars invars);
  FoMo::DataCube readdatacube();
  FoMo::RenderCube readrendering();
  FoMo::GoftCube readgoftcube();
  void setrendermethod(const std::string inrendermethod);
  std::string readrendermethod();
  void setchiantifile(const std::string inchianti);
  void setabundfile(const std::string inabund);
  std::string readchiantifile();
  std::string readabundfile();
  void setoutfile(const std::string outfile);
  void push_back_datapoint(std::vector<double> coordinate, std::vector<double> variables,
                           std::vector<std::string>* unitvec = NULL);
  void setdata(tgrid& ingrid, tvars& indata, std::vector<std::string>* unitvec = NULL);
  void setobservationtype(FoMoObservationType);
  FoMoObservationType readobservationtype();
  void setresolution(const int& x_pixel, const int& y_pixel, const int& z_pixel, const int& lambda_pixel,
                     const double& lambda_width);
  void readresolution(int& x_pixel, int& y_pixel, int& z_pixel, int& lambda_pixel, double& lambda_width);
  void setwriteoptions(std::bitset<noptions> options);
  void setwriteoutbinary(const bool = true);
  void setwriteouttext(const bool = true);
  void setwriteoutzip(const bool = true);
  void setwriteoutdeletefiles(const bool = true);
};

}  // namespace FoMo

#endif
