// gpugi_render: headless, script-driven front end of the HIP renderer.
//
// Reads the reference application's script language (one `name = a, b, c`
// per line, `#` or `//` comments, blanks ignored: ScriptProcessing::ParseCommand,
// control/scriptprocessing.cpp:66-156) from a file or from `-e` arguments and
// drives RendererSystem as Application does (application.cpp:77-260).
// Instead of drawing once per window frame, `waitIterations = n` draws n
// iterations.  Usage:
//   gpugi_render [--device N] [script.txt] [-e 'command'] ...
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "renderer.hpp"

namespace {

using Args = std::vector<std::string>;

struct Command {
  std::string name;
  Args args;
};

// ScriptProcessing::ParseCommand: strip comments and all spaces/tabs, split at
// the first '=' and then at commas.  Returns false for an empty line.
bool ParseCommand(std::string line, Command* out) {
  size_t c = line.find('#');
  if (c == std::string::npos) c = line.find("//");
  if (c != std::string::npos) line.erase(c);
  std::string s;
  for (char ch : line)
    if (ch != ' ' && ch != '\t' && ch != '\r' && ch != '\n') s += ch;
  if (s.empty()) return false;
  const size_t eq = s.find('=');
  out->name = s.substr(0, eq);
  out->args.clear();
  if (eq != std::string::npos) {
    size_t last = eq;
    for (;;) {
      const size_t next = s.find(',', last + 1);
      out->args.push_back(s.substr(last + 1, next == std::string::npos ? std::string::npos : next - last - 1));
      if (next == std::string::npos) break;
      last = next;
    }
  }
  return true;
}

float AsFloat(const std::string& v) {
  char* end = nullptr;
  const float f = std::strtof(v.c_str(), &end);
  if (v.empty() || *end) throw std::invalid_argument("not a number: '" + v + "'");
  return f;
}

int AsInt(const std::string& v) {
  char* end = nullptr;
  const long i = std::strtol(v.c_str(), &end, 10);
  if (v.empty() || *end) throw std::invalid_argument("not an integer: '" + v + "'");
  return (int)i;
}

void NeedArgs(const Command& c, size_t n) {
  if (c.args.size() != n)
    throw std::invalid_argument("'" + c.name + "' expects " + std::to_string(n) + " argument(s)");
}

const char* kHelp =
    "sceneFilename = <file.rawscene> | cornell | heightfield[:cells[:seed]]   load a scene (json beside the rawscene)\n"
    "cameraPos = x, y, z / cameraLookAt = x, y, z / cameraFOV = degrees   pinhole camera\n"
    "resolution = width, height                                           backbuffer size\n"
    "pathLength = n                                                       maximum path length (default 16)\n"
    "renderer = 0 | 1 | 2                  0: Pathtracer, 1: LightPathtracer, 2: BidirectionalPathtracer\n"
    "traversal = 0 | 1 | 2                 integrator traversal: auto, reference escape walk, wide BVH\n"
    "bvhType = 0 | 1                       0: axis-aligned, 1: oriented boxes for the next scene load\n"
    "addPointLight = x, y, z, r, g, b      add a point light to the scene\n"
    "waitIterations = n                    draw n iterations\n"
    "screenshot / namedscreenshot = file   write the image (sum / iterations) as PFM\n"
    "envMap = size, x-, x+, y-, y+, z-, z+  environment cube map from six size x size binary PPM faces\n"
    "referenceImage = file.pfm             reference image of the MSE checks\n"
    "mseCheckInterval = n                  compute and print the MSE every n-th iteration (0: never)\n"
    "waitMSE = x                           draw until the MSE is below x (needs the two above)\n"
    "help                                  this text\n"
    "exit                                  stop processing\n";

class App {
 public:
  explicit App(int device) : m_device(device) {}

  // Returns false on `exit`.
  bool Run(const Command& c) {
    if (c.name == "help") {
      std::cout << kHelp;
    } else if (c.name == "exit") {
      return false;
    } else if (c.name == "bvhType") {
      NeedArgs(c, 1);
      const int t = AsInt(c.args[0]);
      if (t != 0 && t != 1) throw std::invalid_argument("bvhType: 0 (axis-aligned boxes) or 1 (oriented boxes)");
      System().SetBvhType(t);
    } else if (c.name == "sceneFilename") {
      NeedArgs(c, 1);
      LoadScene(c.args[0]);
    } else if (c.name == "cameraPos" || c.name == "cameraLookAt") {
      NeedArgs(c, 3);
      const gpugi::Vec3 v{AsFloat(c.args[0]), AsFloat(c.args[1]), AsFloat(c.args[2])};
      if (c.name == "cameraPos") m_camera.SetPosition(v);
      else m_camera.SetLookAt(v);
      System().SetCamera(m_camera);
    } else if (c.name == "cameraFOV") {
      NeedArgs(c, 1);
      m_camera.SetVFov(AsFloat(c.args[0]));
      System().SetCamera(m_camera);
    } else if (c.name == "resolution") {
      NeedArgs(c, 2);
      const int w = AsInt(c.args[0]), h = AsInt(c.args[1]);
      m_camera.SetAspectRatio((float)w / (float)h);
      System().SetCamera(m_camera);
      System().SetScreenSize(w, h);
    } else if (c.name == "pathLength") {
      NeedArgs(c, 1);
      System().SetPathLength(AsInt(c.args[0]));
    } else if (c.name == "traversal") {
      NeedArgs(c, 1);
      System().SetTraversalKernel(AsInt(c.args[0]));
    } else if (c.name == "renderer") {
      NeedArgs(c, 1);
      SwitchRenderer(AsInt(c.args[0]));
    } else if (c.name == "addPointLight") {
      NeedArgs(c, 6);
      if (!m_scene) throw std::invalid_argument("addPointLight: no scene loaded");
      m_scene->AddPointLight({AsFloat(c.args[0]), AsFloat(c.args[1]), AsFloat(c.args[2])},
                             {AsFloat(c.args[3]), AsFloat(c.args[4]), AsFloat(c.args[5])});
    } else if (c.name == "waitIterations") {
      NeedArgs(c, 1);
      const int n = AsInt(c.args[0]);
      for (int i = 0; i < n; ++i) Draw();
    } else if (c.name == "envMap") {
      NeedArgs(c, 7);
      System().SetEnvironmentMap(AsInt(c.args[0]), c.args[1], c.args[2], c.args[3], c.args[4], c.args[5],
                                 c.args[6]);
    } else if (c.name == "referenceImage") {
      NeedArgs(c, 1);
      Mse().LoadReferenceFromPFM(c.args[0]);
      if (!Mse().HasValidReferenceImage()) std::cerr << "Error reading hdr image from " << c.args[0] << "\n";
    } else if (c.name == "mseCheckInterval") {
      NeedArgs(c, 1);
      m_mseCheckInterval = AsInt(c.args[0]);
    } else if (c.name == "waitMSE") {
      NeedArgs(c, 1);
      WaitMse(AsFloat(c.args[0]));
    } else if (c.name == "screenshot" || c.name == "namedscreenshot") {
      SaveImage(c.name == "namedscreenshot" && !c.args.empty() ? c.args[0] : std::string());
    } else {
      throw std::invalid_argument("unknown parameter '" + c.name + "' (try help)");
    }
    return true;
  }

 private:
  gpugi::RendererSystem& System() {
    if (!m_system) m_system.reset(new gpugi::RendererSystem(m_device));
    return *m_system;
  }

  gpugi::TextureMSE& Mse() {
    if (!m_mse) m_mse.reset(new gpugi::TextureMSE(System()));
    return *m_mse;
  }

  // Application::Draw (application.cpp:328-344): one iteration, then the MSE check.
  void Draw() {
    System().Draw();
    MseChecks();
  }

  // Application::MSEChecks (application.cpp:346-376); returns the MSE when it was computed.
  bool MseChecks(float* mseOut = nullptr) {
    ++m_iterationSinceLastMseCheck;
    if (m_mseCheckInterval <= 0 || m_iterationSinceLastMseCheck < (unsigned)m_mseCheckInterval) return false;
    m_iterationSinceLastMseCheck = 0;
    const float mse = Mse().ComputeCurrentMSE((unsigned)System().GetIterationCount());
    std::cout << "Iteration " << System().GetIterationCount() << " MSE: " << mse << "\n";
    if (mseOut) *mseOut = mse;
    return true;
  }

  // The waitMSE listener (application.cpp:180-188) and the script pause it
  // starts: without a window loop the iterations are drawn here until a check
  // finds the MSE below the goal or NaN (or at once for a goal <= 0).
  void WaitMse(float goal) {
    if (m_mseCheckInterval <= 0) {
      std::cerr << "Can't activate waitMSE as long the mseCheckInterval is 0!\n";
      return;
    }
    if (!Mse().HasValidReferenceImage()) {
      std::cerr << "Can't activate waitMSE as long as there is no reference image loaded!\n";
      return;
    }
    if (!System().GetActiveRenderer()) throw std::invalid_argument("waitMSE: no renderer");
    for (;;) {
      System().Draw();
      float mse = 0.0f;
      if (!MseChecks(&mse)) continue;
      if (!(goal > 0.0f)) return;
      if (mse != mse) {
        std::cerr << "MSE is NaN. Resuming script.\n";
        return;
      }
      if (goal > mse) {
        std::cout << "MSE now below " << goal << "! Resuming script\n";
        return;
      }
    }
  }

  void LoadScene(const std::string& name) {
    if (name == "cornell") {
      m_scene = gpugi::Scene::Cornell();
    } else if (name.rfind("heightfield", 0) == 0) {
      unsigned cells = 64, seed = 0x9E3779B9u;
      std::string rest = name.substr(11);
      if (!rest.empty()) {
        if (rest[0] != ':') throw std::invalid_argument("heightfield[:cells[:seed]]");
        rest = rest.substr(1);
        const size_t k = rest.find(':');
        cells = (unsigned)std::strtoul(rest.substr(0, k).c_str(), nullptr, 0);
        if (k != std::string::npos) seed = (unsigned)std::strtoul(rest.substr(k + 1).c_str(), nullptr, 0);
      }
      m_scene = gpugi::Scene::Heightfield(cells, seed);
    } else {
      std::string json = name;
      const size_t dot = json.rfind('.');
      json = (dot == std::string::npos ? json : json.substr(0, dot)) + ".json";
      std::ifstream probe(json);
      m_scene = gpugi::Scene::Load(name, probe.good() ? json : std::string());
    }
    std::cout << "Loaded scene " << name << " (" << m_scene->GetNumTriangles() << " triangles)\n";
    System().SetScene(m_scene);
  }

  void SwitchRenderer(int index) {
    if (index < 0 || index > 2)
      throw std::invalid_argument("renderer " + std::to_string(index) + " is not part of this build (0, 1, 2)");
    gpugi::RendererSystem& rs = System();
    switch (index) {
      case 0: rs.SetRenderer<gpugi::Pathtracer>(); break;
      case 1: rs.SetRenderer<gpugi::LightPathtracer>(); break;
      default: rs.SetRenderer<gpugi::BidirectionalPathtracer>(); break;
    }
    std::cout << "Switched to " << rs.GetActiveRenderer()->GetName() << "\n";
  }

  // Application::SaveImage (application.cpp:394-414).
  void SaveImage(const std::string& name) {
    gpugi::RendererSystem& rs = System();
    if (!rs.GetActiveRenderer()) return;
    std::string file = name;
    if (file.empty())
      file = rs.GetActiveRenderer()->GetName() + " " + std::to_string(rs.GetIterationCount()) + "it " +
             std::to_string(rs.GetRenderTime()) + "ms.pfm";
    rs.WritePfm(file);
    std::cout << "Wrote screenshot \"" << file << "\" (" << rs.GetIterationCount() << " iterations, "
              << rs.GetRenderTime() << " ms)\n";
  }

  int m_device;
  std::unique_ptr<gpugi::RendererSystem> m_system;
  std::unique_ptr<gpugi::TextureMSE> m_mse;
  int m_mseCheckInterval = 0;
  unsigned m_iterationSinceLastMseCheck = 0;
  std::shared_ptr<gpugi::Scene> m_scene;
  gpugi::Camera m_camera;
};

}  // namespace

int main(int argc, char** argv) {
  int device = 0;
  std::vector<std::string> lines;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--device" && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else if (a == "-e" && i + 1 < argc) {
      lines.push_back(argv[++i]);
    } else if (a == "-h" || a == "--help") {
      std::cout << "usage: gpugi_render [--device N] [script.txt] [-e 'command'] ...\n" << kHelp;
      return 0;
    } else {
      std::ifstream f(a);
      if (!f) {
        std::cerr << "gpugi_render: cannot open script " << a << "\n";
        return 2;
      }
      std::string l;
      while (std::getline(f, l)) lines.push_back(l);
    }
  }
  App app(device);
  int lineNo = 0;
  for (const std::string& l : lines) {
    ++lineNo;
    Command c;
    if (!ParseCommand(l, &c)) continue;
    try {
      if (!app.Run(c)) break;
    } catch (const gpugi::Error& e) {
      std::cerr << "gpugi_render: line " << lineNo << " (" << c.name << "): " << e.what() << "\n";
      return 1;
    } catch (const std::invalid_argument& e) {
      std::cerr << "gpugi_render: line " << lineNo << " (" << c.name << "): " << e.what() << "\n";
      return 2;
    }
  }
  std::cout.flush();
  return 0;
}
