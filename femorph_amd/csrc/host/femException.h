// femException.h -- feMorph's exception type (src/femException.h), for the
// standalone build: constructed from a message, returned by what().
#ifndef FEMEXCEPTION_H
#define FEMEXCEPTION_H

#include <exception>
#include <string>

class femException : public std::exception {
 public:
  femException(const char* m) : msg(m) {}
  virtual ~femException() throw() {}
  virtual const char* what() const throw() { return msg.c_str(); }

 protected:
  std::string msg;
};

#endif  // FEMEXCEPTION_H
