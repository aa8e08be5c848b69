// Test double of libpressio_ext/cpp/pressio.h: the compressor registry and
// the static registration helper plugins use.
#pragma once

#include "compressor.h"

#include <functional>
#include <map>
#include <memory>
#include <string>

template<class Base>
struct pressio_registry
{
  std::map<std::string, std::function<std::unique_ptr<Base>()>> factories;

  template<class F>
  void register_factory(std::string const& name, F factory)
  {
    factories[name] = factory;
  }

  std::unique_ptr<Base> build(std::string const& name) const
  {
    auto it = factories.find(name);
    return it == factories.end() ? nullptr : it->second();
  }

  bool contains(std::string const& name) const { return factories.count(name) != 0; }
};

inline pressio_registry<libpressio_compressor_plugin>&
compressor_plugins()
{
  static pressio_registry<libpressio_compressor_plugin> registry;
  return registry;
}

struct pressio_register
{
  template<class Registry, class Factory>
  pressio_register(Registry& registry, const char* name, Factory factory)
  {
    registry.register_factory(name, factory);
  }
};
